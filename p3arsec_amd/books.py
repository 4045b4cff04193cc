"""Synthetic option books and the PARSEC blackscholes text format.

A *book* is a structure-of-arrays set of European options: spot ``s``,
strike ``k``, risk-free rate ``r``, volatility ``v``, time to expiry ``t``
(years) as float32, and ``otype`` as uint8 (0 = call, 1 = put).  This is the
layout the reference builds from its AoS ``OptionData`` records
(parsec-ff/pkgs/apps/blackscholes/src/blackscholes.c:747-767), with the
option type narrowed from int32 to one byte.

Distributions (seeded, counter-based Philox stream):
    s ~ U[10, 200];  k = s * U[0.5, 1.5];  r ~ U[0.01, 0.10];
    v ~ U[0.05, 0.65];  t ~ U[0.05, 2.0];  type ~ Bernoulli(0.5); q = 0.
Seeds are 20261202 + workload id (1 = 4096-option book, 2 = 10M book,
3 = FP64 book, 4 = Greeks book, 5 = serving books).
"""
from __future__ import annotations

import dataclasses
import numpy as np

SEED_BASE = 20261202
FIELDS = ("s", "k", "r", "v", "t")


@dataclasses.dataclass
class Book:
    s: np.ndarray
    k: np.ndarray
    r: np.ndarray
    v: np.ndarray
    t: np.ndarray
    otype: np.ndarray

    @property
    def n(self) -> int:
        return int(self.s.shape[0])

    def arrays(self):
        return (self.s, self.k, self.r, self.v, self.t, self.otype)

    def astype(self, dtype) -> "Book":
        return Book(*(a.astype(dtype) for a in (self.s, self.k, self.r, self.v, self.t)),
                    self.otype.copy())

    def slice(self, lo: int, hi: int) -> "Book":
        return Book(*(a[lo:hi] for a in self.arrays()))


def make_book(n: int, seed: int, dtype=np.float32) -> Book:
    """Seeded synthetic book with the distributions above."""
    g = np.random.Generator(np.random.Philox(seed))
    s = g.uniform(10.0, 200.0, n)
    k = s * g.uniform(0.5, 1.5, n)
    r = g.uniform(0.01, 0.10, n)
    v = g.uniform(0.05, 0.65, n)
    t = g.uniform(0.05, 2.0, n)
    otype = (g.random(n) < 0.5).astype(np.uint8)
    return Book(s.astype(dtype), k.astype(dtype), r.astype(dtype), v.astype(dtype),
                t.astype(dtype), otype)


def edge_book() -> Book:
    """Edge cases at and beyond the corners of the distributions: S == K,
    S/K one ulp off 1, the minimum sigma*sqrt(T) (0.05 * sqrt(0.05)), T at
    0.05 and 2.0, K/S at 0.5 and 1.5, r and sigma at their bounds, deep in
    and out of the money, and the Hull textbook cases; every row as call and
    as put."""
    rows = []
    f32 = np.float32
    for s in (10.0, 42.0, 100.0, 200.0):
        for ks in (0.5, 1.0, 1.5):
            for r in (0.01, 0.10):
                for v in (0.05, 0.65):
                    for t in (0.05, 2.0):
                        rows.append((s, s * ks, r, v, t))
    for s in (f32(10.0), f32(97.5), f32(200.0)):
        up = np.nextafter(s, f32(1e9))
        dn = np.nextafter(s, f32(0.0))
        for k in (up, dn, s):
            rows.append((float(s), float(k), 0.05, 0.05, 0.05))
            rows.append((float(s), float(k), 0.01, 0.65, 2.0))
    rows += [
        (42.0, 40.0, 0.1, 0.2, 0.5),       # Hull: call 4.7594, put 0.8086
        (100.0, 100.0, 0.05, 0.3, 1.0),    # call 14.2312, put 9.3542
        (200.0, 10.0, 0.1, 0.05, 2.0),     # deep ITM call / deep OTM put
        (10.0, 200.0, 0.01, 0.05, 0.05),   # deep OTM call / deep ITM put
        (150.0, 75.0, 0.10, 0.05, 0.05),
        (75.0, 112.5, 0.01, 0.65, 2.0),
    ]
    a = np.array(rows, dtype=np.float64)
    n = a.shape[0]
    cols = [np.concatenate([a[:, j], a[:, j]]).astype(np.float32) for j in range(5)]
    otype = np.concatenate([np.zeros(n, np.uint8), np.ones(n, np.uint8)])
    return Book(*cols, otype)


def degenerate_book() -> Book:
    """Inputs outside the distributions, where the reference's arithmetic
    meets IEEE special values: T = 0 and sigma = 0 (division by zero, d =
    +-inf), S = K with T = 0 (0/0), S = 0 and K = 0 (log of 0 and of inf),
    negative T (square root of a negative), zero and negative rates, very
    large and very small magnitudes, and NaN and infinite inputs; every row
    as call and as put."""
    inf, nan = float("inf"), float("nan")
    rows = [
        (100.0, 90.0, 0.05, 0.3, 0.0), (90.0, 100.0, 0.05, 0.3, 0.0), (100.0, 100.0, 0.05, 0.3, 0.0),
        (100.0, 90.0, 0.05, 0.0, 1.0), (90.0, 100.0, 0.05, 0.0, 1.0), (100.0, 100.0, 0.0, 0.0, 1.0),
        (100.0, 90.0, 0.05, 0.0, 0.0), (0.0, 100.0, 0.05, 0.3, 1.0), (100.0, 0.0, 0.05, 0.3, 1.0),
        (0.0, 0.0, 0.05, 0.3, 1.0), (100.0, 90.0, 0.05, 0.3, -1.0), (100.0, 90.0, 0.0, 0.3, 1.0),
        (100.0, 90.0, -0.05, 0.3, 1.0), (1e30, 1e30, 0.05, 0.3, 1.0), (1e-30, 1e-30, 0.05, 0.3, 1.0),
        (100.0, 90.0, 0.05, 1e-30, 1.0), (100.0, 90.0, 0.05, 1e10, 1.0), (100.0, 90.0, 0.05, 0.3, 1e-30),
        (100.0, 90.0, 0.05, 0.3, 1e10), (100.0, 90.0, 50.0, 0.3, 1.0), (inf, 100.0, 0.05, 0.3, 1.0),
        (100.0, inf, 0.05, 0.3, 1.0), (100.0, 90.0, inf, 0.3, 1.0), (100.0, 90.0, 0.05, inf, 1.0),
        (100.0, 90.0, 0.05, 0.3, inf), (nan, 100.0, 0.05, 0.3, 1.0), (100.0, nan, 0.05, 0.3, 1.0),
        (100.0, 90.0, nan, 0.3, 1.0), (100.0, 90.0, 0.05, nan, 1.0), (100.0, 90.0, 0.05, 0.3, nan),
        (-100.0, 90.0, 0.05, 0.3, 1.0), (100.0, -90.0, 0.05, 0.3, 1.0),
    ]
    a = np.array(rows, dtype=np.float64)
    n = a.shape[0]
    cols = [np.concatenate([a[:, j], a[:, j]]).astype(np.float32) for j in range(5)]
    otype = np.concatenate([np.zeros(n, np.uint8), np.ones(n, np.uint8)])
    return Book(*cols, otype)


# ---------------------------------------------------------------------------
# PARSEC text format (blackscholes.c:701-734 read, :923-947 write).
def write_parsec_input(path: str, book: Book) -> None:
    """`N` then one line per option: "s strike r divq v t type divs DGrefval".
    Floats are written with 9 significant digits so that fscanf("%f") reads
    back the exact float32; divq, divs and DGrefval are 0."""
    with open(path, "w") as f:
        f.write(f"{book.n}\n")
        tchar = np.where(book.otype != 0, "P", "C")
        for i in range(book.n):
            f.write("%.9g %.9g %.9g 0 %.9g %.9g %s 0 0\n" % (
                book.s[i], book.k[i], book.r[i], book.v[i], book.t[i], tchar[i]))


def read_parsec_output(path: str) -> np.ndarray:
    """Prices written as "%i" then N x "%.18f" (blackscholes.c:930-937)."""
    with open(path) as f:
        n = int(f.readline())
        vals = np.array([float(x) for x in f.read().split()], dtype=np.float64)
    if vals.shape[0] != n:
        raise ValueError(f"{path}: header says {n} prices, found {vals.shape[0]}")
    return vals.astype(np.float32)
