"""CPU tests of the swaptions oracle (oracle/swaptions_oracle.c).

The routines HJM_Swaption_Blocking calls are not in the reference tree
(PARSEC 3.0's RanUnif.cpp, CumNormalInv.cpp, HJM.cpp,
HJM_SimPath_Forward_Blocking.cpp), so the oracle restates their published
algorithms and these tests pin each one against what the mathematics fixes:
Park-Miller's published check value, Schrage's factorisation against direct
modular arithmetic, Moro against the exact inverse normal, the forward-curve
and drift identities, deterministic and deep in-the-money prices against
their closed forms (the discounted bond prices are martingales under the HJM
drift), and the per-chain form the GPU kernel uses against the blocked form.
"""
import math
import os
import sys

import numpy as np
import pytest
from scipy.special import ndtri

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

import pyswaptions as S  # noqa: E402
from p3arsec_amd import swaptions as W  # noqa: E402

YLD = W.parsec_portfolio(1)[0][0]["yield"]


def sw(**kw):
    d = {"strike": 0.1, "compounding": 0.0, "maturity": 1.0, "tenor": 2.0, "interval": 1.0,
         "years": 5.5, "yield": YLD, "factors": W.PARSEC_FACTORS}
    d.update(kw)
    return d


def test_minstd_published_check_value():
    # Park & Miller (CACM 1988): 10000 steps of x <- 16807 x mod (2^31-1) from 1
    assert S.minstd_iterate(1, 10000) == 1043618065


def test_schrage_equals_direct_modular_arithmetic():
    rng = np.random.default_rng(5)
    m = 2147483647
    cs = list(rng.integers(0, 2 ** 42, 5000)) + [0, 1, m - 1, m, m + 1, 2 * m, 2 ** 42 - 1]
    for c in cs:
        assert S.ranunif(int(c)) == S.ranunif_direct(int(c))
    # the draw is MINSTD of the hashed counter, scaled by 4.656612875e-10
    c = 123456
    x = (16807 * ((c * 1513517) % m)) % m
    assert S.ranunif(c) == x * 4.656612875e-10


def test_host_generator_matches_oracle():
    for seed in (1979, 0, 7, 2 ** 31 + 5, -1, -1979, -(2 ** 31) - 3):
        sws, base = W.parsec_portfolio(40, seed)
        b2, years, strike, _ = S.portfolio(40, seed)
        assert base == b2
        assert np.array_equal([s["years"] for s in sws], years)
        assert np.array_equal([s["strike"] for s in sws], strike)
    # the driver's ranges: horizons 5..19.75 in quarters, strikes 0.1..4.9
    sws, _ = W.parsec_portfolio(500)
    y = np.array([s["years"] for s in sws])
    k = np.array([s["strike"] for s in sws])
    assert y.min() >= 5.0 and y.max() <= 19.75 and np.all((y * 4) == np.round(y * 4))
    assert k.min() >= 0.1 - 1e-12 and k.max() <= 4.9 + 1e-12


def test_moro_against_exact_inverse_normal():
    u = np.concatenate([np.linspace(1e-10, 1 - 1e-10, 40001), [0.08, 0.92, 0.0800001, 0.5]])
    err = np.abs(S.cum_normal_inv(u) - ndtri(u))
    # Moro's published accuracy: about 3e-9 over the central and tail regions
    assert err.max() < 3.5e-9
    z = S.cum_normal_inv(np.array([0.25, 0.75]))   # u - 0.5 = -+0.25 exactly
    assert z[0] == -z[1]


def test_forward_curve_identity():
    rng = np.random.default_rng(1)
    for _ in range(20):
        y = rng.uniform(0.0, 0.2, 12)
        f = S.yield_to_forward(y)
        # the average of the forwards up to i is the yield at i
        for i in range(len(y)):
            assert math.isclose(f[: i + 1].sum(), (i + 1) * y[i], rel_tol=1e-13, abs_tol=1e-15)


def test_drift_condition():
    rng = np.random.default_rng(2)
    for F, n, years in ((3, 11, 5.5), (1, 5, 2.0), (5, 20, 17.25)):
        fa = rng.uniform(-0.01, 0.02, (F, n - 1))
        total, per = S.drifts(years, fa)
        dt = years / n
        for k in range(F):
            cs = np.cumsum(per[k])
            want = 0.5 * dt * np.cumsum(fa[k]) ** 2
            np.testing.assert_allclose(cs, want, rtol=1e-12, atol=1e-18)
        np.testing.assert_allclose(total, per.sum(axis=0), rtol=1e-15, atol=0)


SHAPES = [
    dict(),
    dict(years=19.75, strike=0.3),
    dict(years=5.0, strike=0.2, compounding=0.5),
    dict(years=8.0, maturity=2.0, tenor=3.0, interval=0.5, strike=0.12),
    dict(years=10.0, maturity=0.0, tenor=4.0, strike=0.15),          # starts at t = 0
    dict(years=3.0, maturity=2.1, tenor=0.5, interval=0.25, strike=0.14),   # curve past horizon
    dict(years=6.0, factors=W.PARSEC_FACTORS[:1], strike=0.13),
    dict(years=12.0, n=20, F=5, maturity=4.0, tenor=5.0, strike=0.14),
]


def shape(d):
    d = dict(d)
    n = d.pop("n", None)
    F = d.pop("F", None)
    if n is not None:
        rng = np.random.default_rng(n * 10 + F)
        d["yield"] = 0.05 + 0.004 * np.arange(n)
        d["factors"] = rng.uniform(0.0, 0.01, (F, n - 1))
    return sw(**d)


@pytest.mark.parametrize("i", range(len(SHAPES)))
def test_chain_form_equals_blocked_form(i):
    s = shape(SHAPES[i])
    for trials, block in ((1000, 16), (1001, 16), (77, 1), (500, 7)):
        a = S.price_blocking(s, 1234567 + i, trials, block)
        b = S.price_chain(s, 1234567 + i, trials, block)
        assert a == b or (np.isnan(a[1]) and np.isnan(b[1]) and a[0] == b[0])


def discount(f0, dt, j):
    return math.exp(-dt * sum(f0[:j]))


def closed_form(s):
    """Price of a swaption in the money on every path: sum_i c_i P(0, T_m + i) -
    P(0, T_m), for a swap inside the simulated horizon."""
    n = len(s["yield"])
    dt = s["years"] / n
    m = int(s["maturity"] / dt + 0.5)
    L = int(n - s["maturity"] / dt + 0.5)
    freq = int(s["interval"] / dt + 0.5)
    pts = int(s["tenor"] / dt + 0.5)
    k = s["strike"] if s["compounding"] == 0 else math.log(1 + s["strike"] * s["compounding"]) / s["compounding"]
    f0 = S.yield_to_forward(s["yield"])
    assert m + L <= n + 1
    v = -discount(f0, dt, m)
    for i in range(freq, pts + 1, freq):
        c = math.exp(k * s["interval"]) - (1 if i != pts else 0)
        v += c * discount(f0, dt, m + i)
    return v


# The model apart from the generator: independent normal draws, four times
# the driver's volatilities and 200000 paths.  Dropping the HJM drift then
# moves these prices by 15 standard errors or more.
@pytest.mark.parametrize("kw", [dict(years=5.5, strike=2.0), dict(years=14.0, strike=0.9),
                                dict(years=8.0, maturity=2.0, tenor=3.0, interval=0.5,
                                     strike=1.5), dict(years=10.0, maturity=0.0, strike=0.8)])
def test_deep_in_the_money_price_is_the_forward_swap_value(kw):
    s = sw(factors=4 * W.PARSEC_FACTORS, **kw)
    m = S.start_step(s)
    z = np.random.default_rng(11).standard_normal((200000, max(m, 1) * 3))
    v = S.chain_payoffs(s, z)
    assert v.min() > 0          # in the money on every path
    p, se = v.mean(), v.std(ddof=1) / math.sqrt(len(v))
    cf = closed_form(s)
    assert abs(p - cf) < 5 * se + 1e-12, (p, cf, se)


def test_generator_is_a_weyl_sequence_in_the_counter():
    # u(c) = (a c mod M) / M with a = 16807 * 1513517 mod M: consecutive
    # counters step by a fixed amount mod 1 (why PARSEC's estimates carry a
    # bias of their own, which the map reproduces rather than corrects)
    m = 2147483647
    a = 16807 * 1513517 % m
    for c in (0, 1, 999, 2 ** 35 + 17):
        x0 = round(S.ranunif(c) / 4.656612875e-10)
        x1 = round(S.ranunif(c + 1) / 4.656612875e-10)
        assert (x1 - x0) % m == a


def test_zero_volatility_price_is_deterministic():
    s = sw(years=7.0, strike=0.16, factors=np.zeros((3, 10)))
    p, _ = S.price_blocking(s, 42, 64, 16)
    n = 11
    dt = s["years"] / n
    f0 = S.yield_to_forward(s["yield"])
    m = int(1.0 / dt + 0.5)
    # with no volatility the curve at T_m is the t = 0 curve shifted by m
    assert abs(p - max(closed_form(s), 0.0)) < 1e-12 * max(1.0, abs(p))
    assert m == 2


def test_static_blocks_thread_invariance():
    sws, base = W.parsec_portfolio(9)
    a = S.price_batch(sws, base, 2000, 16, nthreads=1)
    b = S.price_batch(sws, base, 2000, 16, nthreads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # the pruned per-chain form (the CPU baseline's second variant): same bits
    c = S.price_batch(sws, base, 2000, 16, nthreads=4, chain=True)
    assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])
    for i in (0, 8):
        assert S.price_blocking(sws[i], base + i, 2000) == (a[0][i], a[1][i])


def test_unusable_shapes_are_refused():
    with pytest.raises(ValueError):
        S.price_blocking(sw(maturity=6.0), 1, 100)       # start past the horizon
    with pytest.raises(ValueError):
        S.price_blocking(sw(interval=0.1), 1, 100)       # payments under one step
    with pytest.raises(ValueError):
        S.price_blocking(sw(tenor=9.0), 1, 100)          # swap past the curve
