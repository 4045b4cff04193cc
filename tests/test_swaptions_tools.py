"""The swaptions drop-ins: the PARSEC `swaptions` command line
(p3arsec_amd/bin/swaptions, the reference driver HJM_Securities.cpp:153-378)
and bs_compat::HJM_Swaption_Blocking (HJM_Swaption_Blocking.cpp:20-40).
Argument handling runs on any host; pricing needs the GPU and is checked
against the oracle's restatement of HJM_Swaption_Blocking."""
import math
import os
import re
import subprocess

import pytest

from conftest import ROOT

CLI = os.path.join(ROOT, "p3arsec_amd", "bin", "swaptions")
DEMO = os.path.join(ROOT, "tests", "native", "swaption_compat_demo")
LINE = re.compile(r"Swaption (\d+): \[SwaptionPrice: (\S+) StdError: (\S+)\]")


def run(*args):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=120)


def test_cli_arguments():
    assert os.path.exists(CLI), "build the tools: make -C p3arsec_amd/csrc tools"
    p = run()
    assert p.returncode == 1 and "Usage" in p.stderr and "PARSEC Benchmark Suite" in p.stdout
    p = run("-ns", 2, "-nt", 4)
    assert p.returncode == 1 and "Fewer swaptions than threads" in p.stderr
    p = run("-ns", 4, "-xx", 1)
    assert p.returncode == 1 and "Unknown option: -xx" in p.stderr
    p = run("-ns", 2000, "-nt", 2000)
    assert p.returncode == 1 and "between 1 and 1024" in p.stderr


def close(a, b):
    return abs(a - b) <= 1e-12 * abs(b) + 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1979, 12345, -77])
def test_cli_matches_oracle(gpu, seed):
    import pyswaptions as S
    from p3arsec_amd import swaptions as W
    p = run("-ns", 12, "-sm", 20000, "-nt", 4, "-sd", seed)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Number of Simulations: 20000,  Number of threads: 4 Number of swaptions: 12" in p.stdout
    assert re.search(r"roi\.time\|[0-9.]+", p.stdout)
    rows = [(int(i), float(a), float(b)) for i, a, b in LINE.findall(p.stderr)]
    assert [r[0] for r in rows] == list(range(12))
    sws, base = W.parsec_portfolio(12, seed)
    for i, price, se in rows:
        op, oe = S.price_blocking(sws[i], base + i, 20000)
        # the CLI prints 10 decimals
        assert abs(price - op) <= 1e-10 + 1e-12 * abs(op), (i, price, op)
        assert abs(se - oe) <= 1e-10 + 1e-8 * oe, (i, se, oe)


@pytest.mark.gpu
def test_compat_hjm_swaption_blocking(gpu):
    import pyswaptions as S
    from p3arsec_amd import swaptions as W
    for years, strike, seed in ((5.5, 0.13, 7), (16.5, 3.1, 1000), (9.25, 0.1, 2 ** 31)):
        p = subprocess.run([DEMO, str(years), str(strike), str(seed), "3001"],
                           capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        ok, price, se = p.stdout.split()
        sw = W.parsec_portfolio(1)[0][0]
        sw.update(years=years, strike=strike)
        op, oe = S.price_blocking(sw, seed, 3001)
        assert ok == "1" and close(float(price), op)
        # 3001 trials simulate 3008 paths but divide by 3001: for a swaption deep in
        # the money sumsq - sum^2/3001 is negative and the reference's error is NaN
        se = float(se)
        assert (math.isnan(se) and math.isnan(oe)) or abs(se - oe) <= 1e-8 * oe + 1e-15, (se, oe)
    p = subprocess.run([DEMO, "5.5", "0.1", "-1", "100"], capture_output=True, text=True)
    assert p.returncode == 1 and "seed" in p.stderr       # refused, as an exception


@pytest.mark.gpu
def test_compat_price_swaptions_over_parm_records(gpu):
    """bs_compat::price_swaptions over a record with the reference's parm field
    names: the whole driver loop in one call, each swaption as the oracle."""
    import pyswaptions as S
    from p3arsec_amd import swaptions as W
    p = subprocess.run([DEMO, "map", "24", "5000"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    rows = [line.split() for line in p.stdout.strip().splitlines()]
    sws, base = W.parsec_portfolio(24)
    assert len(rows) == 24
    for i, price, se in rows:
        i = int(i)
        op, oe = S.price_blocking(sws[i], base + i, 5000)
        assert close(float(price), op), (i, price, op)
        se = float(se)
        assert (math.isnan(se) and math.isnan(oe)) or abs(se - oe) <= 1e-8 * oe + 1e-15, (se, oe)
