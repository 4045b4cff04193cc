"""Swaptions (bs_swaptions_host_f64) on the MI355X, through the C ABI, against
the oracle's restatement of HJM_Swaption_Blocking
(parsec-ff/pkgs/apps/swaptions/src/HJM_Swaption_Blocking.cpp:20-222).

The kernel follows the reference's arithmetic step for step (same draws,
same order of every FP64 product and sum); what differs is the device's
exp/log (last-bit differences) and the order of the final sums over paths.
The bar: every price within 1e-12 relative (+1e-13 absolute) of the oracle's
and every standard error within 1e-8 relative, whose formula subtracts two
nearly equal sums (below a noise floor of 8 |price| sqrt(1e-15), where the
payoffs do not vary, both standard errors are rounding noise).  Shapes cover the driver's portfolio, every factor count
1..8, swaptions starting at t = 0 and past the kernel's register form (the
generic form), swap curves past the horizon, compounded strikes, trial
counts that are not multiples of the block size, and the full 128-swaption
1,000,000-trial workload (two swaptions checked against the oracle).
"""
import ctypes
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyswaptions as S  # noqa: E402
from p3arsec_amd import swaptions as W  # noqa: E402

pytestmark = pytest.mark.gpu

YLD = W.parsec_portfolio(1)[0][0]["yield"]


def sw(**kw):
    d = {"strike": 0.1, "compounding": 0.0, "maturity": 1.0, "tenor": 2.0, "interval": 1.0,
         "years": 5.5, "yield": YLD, "factors": W.PARSEC_FACTORS}
    n = kw.pop("n", None)
    F = kw.pop("F", None)
    if n is not None:
        rng = np.random.default_rng(n * 10 + F)
        d["yield"] = 0.05 + 0.004 * np.arange(n)
        d["factors"] = rng.uniform(0.0, 0.01, (F, n - 1))
    d.update(kw)
    return d


def check(got, want, tag):
    (gp, ge), (op, oe) = got, want
    if math.isnan(op) or math.isinf(op):
        assert gp == op or (math.isnan(gp) and math.isnan(op)), (tag, gp, op)
        return 0.0
    assert abs(gp - op) <= 1e-12 * abs(op) + 1e-13, (tag, gp, op)
    # When the payoffs barely vary (a swaption starting at t = 0 pays the same
    # on every path), sumsq - sum^2/n is rounding noise in both, its sign and
    # size set by the order of the sums: at or below this floor (or NaN, the
    # root of a negative noise) the two standard errors are both noise.
    floor = 8 * abs(op) * math.sqrt(1e-15)
    if math.isnan(oe) or oe <= floor:
        assert math.isnan(ge) or ge <= floor, (tag, ge, oe, floor)
    else:
        assert abs(ge - oe) <= 1e-8 * abs(oe) + 1e-15, (tag, ge, oe)
    return abs(gp - op) / max(abs(op), 1e-300)


def test_parsec_portfolio_matches_oracle(gpu):
    sws, base = W.parsec_portfolio(16)
    p, e = W.price(sws, 20000, base)
    worst = 0.0
    for i, s in enumerate(sws):
        worst = max(worst, check((p[i], e[i]), S.price_blocking(s, base + i, 20000), i))
    print(f"16 swaptions x 20000 trials: max relative price difference {worst:.2e}")
    assert np.all(np.isfinite(p)) and np.all(p >= 0)


SHAPES = [   # strikes near the money: 20-40% of paths pay
    dict(strike=0.13),
    dict(years=19.75, strike=0.18),
    dict(years=5.0, strike=0.12, compounding=0.5),
    dict(years=8.0, maturity=2.0, tenor=3.0, interval=0.5, strike=0.19),
    dict(years=10.0, maturity=0.0, tenor=4.0, strike=0.15),               # starts at t = 0
    dict(years=3.0, maturity=2.1, tenor=0.5, interval=0.25, strike=0.19),  # curve past horizon
    dict(years=6.0, factors=W.PARSEC_FACTORS[:1], strike=0.14),            # 1 factor
    dict(n=20, F=5, years=12.0, maturity=4.0, tenor=5.0, strike=0.14),
    dict(n=30, F=2, years=15.0, maturity=3.0, tenor=4.0, strike=0.12),     # start 6: generic
    dict(n=12, F=8, years=6.0, maturity=1.5, tenor=2.0, strike=0.08),      # 8 factors
    dict(n=20, F=8, years=10.0, maturity=4.0, tenor=2.0, strike=0.11),     # 8 x 8 draws
    dict(n=64, F=1, years=16.0, maturity=10.0, tenor=5.0, strike=0.41),    # 64 steps, 40 draws
    dict(n=2, F=3, years=2.0, maturity=0.0, tenor=1.0, strike=0.3,
         factors=W.PARSEC_FACTORS[:, :1]),                                 # smallest curve
    # strongly negative rates: the discount factors past the last cash flow
    # grow, so every path takes the kernel's full computation
    dict(years=11.0, tenor=2.0, strike=-0.3, **{"yield": np.full(11, -0.4)}),
]


@pytest.mark.parametrize("i", range(len(SHAPES)))
def test_shapes_match_oracle(gpu, i):
    s = sw(**SHAPES[i])
    for trials, block in ((1001, 16), (77, 1), (500, 7), (4096, 16)):
        got = W.price([s], trials, 4321 + i, block)
        check((got[0][0], got[1][0]), S.price_blocking(s, 4321 + i, trials, block),
              (i, trials, block))


def test_many_swaptions_mixed_starts(gpu):
    # one launch for swaptions of different horizons, starts and curve lengths
    sws = [sw(years=y, strike=k, maturity=m) for y, k, m in
           ((5.0, 0.11, 1.0), (9.0, 0.13, 2.5), (14.0, 0.2, 3.0), (6.0, 0.1, 0.0),
            (7.5, 0.12, 1.4))] * 7
    p, e = W.price(sws, 3000, 777)
    for i, s in enumerate(sws):
        check((p[i], e[i]), S.price_blocking(s, 777 + i, 3000), i)


def test_bit_reproducible_and_stream(gpu):
    sws, base = W.parsec_portfolio(8)
    a = W.price(sws, 50000, base)
    b = W.price(sws, 50000, base)
    st = gpu.Stream()
    c = W.price(sws, 50000, base, stream=st.handle)
    st.destroy()
    for x, y in ((a, b), (a, c)):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


def test_one_trial_and_single_path_blocks(gpu):
    s = sw(years=5.5, strike=0.3)
    for trials, block in ((1, 1), (1, 16), (2, 1)):
        got = W.price([s], trials, 99, block)
        check((got[0][0], got[1][0]), S.price_blocking(s, 99, trials, block), (trials, block))


def test_zero_draw_gives_the_reference_nan_and_inf(gpu):
    """Counter 2^31 - 1 draws u = 0, whose inverse normal is -inf: that path's
    rates become infinite or NaN (0 x inf where a factor is 0).  The map
    follows the reference through it: same price, NaN or inf included."""
    sws, _ = W.parsec_portfolio(4)
    m31 = 2147483647
    for i, s in enumerate(sws):
        for back in (1, 5, 29, 31, 95):
            seed = m31 - back
            got = W.price([s], 64, seed)
            check((got[0][0], got[1][0]), S.price_blocking(s, seed, 64), (i, back))


def test_native_workload(gpu):
    """PARSEC's native run: 128 swaptions x 1,000,000 trials."""
    sws, base = W.parsec_portfolio(128)
    p, e = W.price(sws, 1_000_000, base)
    assert np.all(np.isfinite(p)) and np.all(p >= 0)
    assert np.all(e[p > 0] > 0)
    for i in (0, 127):
        check((p[i], e[i]), S.price_chain(sws[i], base + i, 1_000_000), i)
    print(f"128 x 1M: prices {p.min():.4f} .. {p.max():.4f}")


def test_counters_near_the_limit(gpu):
    """Seeds whose last counters come within a few thousand of 2^42, where
    1513517 c approaches 2^63: the kernel's Mersenne reduction against the
    oracle's literal 64-bit remainders."""
    s = sw(years=5.5, strike=0.13)
    trials = 1000                      # 1008 paths x 30 counters each
    for slack in (5, 1000, 123457):
        seed = 2 ** 42 - 1008 * 30 - slack
        got = W.price([s], trials, seed)
        check((got[0][0], got[1][0]), S.price_blocking(s, seed, trials), slack)
    # one past the limit is refused
    from p3arsec_amd import capi
    with pytest.raises(capi.BSError):
        W.price([s], trials, 2 ** 42 - 1008 * 30 + 1)
