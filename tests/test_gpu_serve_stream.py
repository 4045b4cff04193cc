"""The serving loop under its configured stream, results checked.

bs_serve_bench (a C++ client of the C ABI) registers a catalog of 256 books
whose sizes are log-uniform on [1, 2^20] options (the batch benchmark's option
distributions), prices every book alone with the batch entry points
(bs_price_f32, bs_price_greeks_f32), then replays a seeded Poisson stream of
requests, 70% price and 30% price + Greeks, against a 32 GB arena.  Every
K-th completed request's outputs are copied back and compared bit for bit
with that book's outputs priced alone.  Streams: the configured 20K req/s
with direct launches (the default) and with graph replay, and 1M req/s,
where batches fuse many requests and their kernels overlap on the device.
"""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "p3arsec_amd", "bin", "bs_serve_bench")


def run_stream(*extra):
    assert os.access(EXE, os.X_OK), f"{EXE} is not built"
    cmd = [EXE, "--budget-gb", "32", "--books", "256", "--max-book", str(1 << 20),
           "--greeks-frac", "0.3"] + [str(x) for x in extra]
    print("running", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr[-2000:]
    r = json.loads(p.stdout.strip().splitlines()[-1])
    print({k: r[k] for k in ("requests", "options_per_s", "p50_ms", "p99_ms", "p99_ms_le64k",
                             "mean_requests_per_launch", "verified", "mismatched")}, flush=True)
    return r


def check(r, min_requests):
    assert r["mismatched"] == 0
    v = r["verified"]
    for key in ("price_le64k", "price_gt64k", "greeks_le64k", "greeks_gt64k"):
        assert v[key] > 0, f"no {key} request verified"
    assert r["requests"] >= min_requests
    assert r["admission_waits"] >= 0


def test_poisson_20k_direct_every_other_result_equals_alone():
    r = run_stream("--rate", 20000, "--duration-s", 3, "--verify-every", 2, "--seed", 11)
    check(r, 0.9 * 20000 * 3)


def test_poisson_20k_graph_replay_results_equal_alone():
    r = run_stream("--rate", 20000, "--duration-s", 2, "--verify-every", 2, "--use-graph", 1,
                   "--seed", 12)
    check(r, 0.9 * 20000 * 2)


def test_poisson_1m_fused_batches_results_equal_alone():
    r = run_stream("--rate", 1000000, "--duration-s", 1.5, "--verify-every", 256, "--seed", 13)
    check(r, 100000)
    assert r["mean_requests_per_launch"] > 2.0   # the mix: requests share batches
