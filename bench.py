"""Benchmark of the Black-Scholes map on one MI355X.

    python bench.py [--gpus 1] [--steps 100] [--warmup 10]

A step is one pass of the price map over the 10,000,000-option FP32 book
(SoA: s, k, r, v, t float32 + uint8 type in, float32 price out; 25 bytes per
option), one kernel launch per pass, as the reference's NUM_RUNS loop
(blackscholes.c:316-347).  Inputs are resident in HBM before timing starts.
Passes rotate over --books device copies of the book (4 x 250 MB), so the
256 MB MALL cannot hold the working set between passes, and alternate
between two HIP streams, so the last waves of one pass overlap the first
waves of the next (passes are independent: adjacent passes use different
books); the same passes on one stream are reported beside it.

Prints one JSON line: options/s over the timed passes, the roofline of the
price kernel (algorithmic bytes / average launch time vs 8 TB/s), and the
CPU baseline (the oracle's FastFlow-style static-block port on the host's
cores, bounded sample).  Extra sections: the fused Greeks kernel's time ratio
to price-only and the FP64 map.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from p3arsec_amd import books, capi  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = ("options/sec & % HBM roofline (10M fp32); Greeks cost ratio; serve p50/p99 ms")
ARGS = None


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def profile_record(kernel: str, n: int) -> dict:
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)[kernel]
    except (OSError, KeyError, ValueError):
        return {}
    return rec if rec.get("options") == n else {}


def measured_traffic(kernel: str, n: int):
    """HBM bytes per launch of `kernel` on the n-option book, from the committed
    rocprofv3 counter runs (profiles/hbm_traffic.json: FETCH_SIZE x 2 +
    WRITE_SIZE x 1, factors calibrated on a plain 256 MiB copy).  None when
    no profile of this size is present."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)[kernel]
    except (OSError, KeyError, ValueError):
        return None
    if rec.get("options") != n:
        return None
    return rec["traffic_bytes"]


LAUNCH_CPU_US = {}
# Streams the passes of each map alternate over, the faster schedule measured
# on MI355X (10M options): price 37.9 us per pass on two streams vs 40.5 on
# one; FP64 109.5 vs 113; the write-heavy Greeks map 70.5 us on one stream vs
# 73.3 on two.
STREAMS = {"price": 2, "fast": 2, "greeks": 1, "fp64": 2}


def section_failed(result: dict, name: str, e: Exception) -> None:
    """Records a failed optional section in the JSON line and on stderr."""
    import traceback
    traceback.print_exc()
    result.setdefault("errors", {})[name] = f"{type(e).__name__}: {e}"[:500]


def spin_up(seconds: float = 0.25) -> None:
    """Brings the GPU to its working clocks before anything is timed: plain
    256 MiB device-to-device copies for `seconds` (the clocks need about 30 ms
    of load to settle; a timed region that starts cold reads 10-20% slow, and
    so does a profiler's per-kernel average over a cold warm-up)."""
    nbytes = 256 << 20
    src, dst = capi.DeviceBuffer(nbytes), capi.DeviceBuffer(nbytes)
    t0 = time.time()
    while time.time() - t0 < seconds:
        for _ in range(20):
            capi.call("bs_memcpy_d2d", dst.ptr, src.ptr, nbytes, None)
        capi.call("bs_device_sync")
    src.free()
    dst.free()


def timed_passes(launch, nbooks: int, steps: int, warmup: int, streams, tag: str = "") -> float:
    """Runs warmup + steps passes, pass i on book i % nbooks and stream
    i % len(streams) (launch(i, stream)); returns ms per pass of the timed
    passes.  The timed region is bracketed by HIP events on the first stream:
    the other streams wait for the start event, and the end event waits for
    every stream's last pass.  Passes on different streams touch different
    books and outputs (nbooks is a multiple of the stream count), so with two
    streams one pass's last waves overlap the next pass's first ones.  Also
    records the host time per launch call (LAUNCH_CPU_US[tag])."""
    assert nbooks % len(streams) == 0
    for i in range(warmup):
        launch(i % nbooks, streams[i % len(streams)].handle)
    capi.call("bs_device_sync")
    e0, e1 = capi.Event(), capi.Event()
    joins = [capi.Event() for _ in streams[1:]]
    e0.record(streams[0].handle)
    for s in streams[1:]:
        s.wait(e0)
    t0 = time.perf_counter()
    for i in range(steps):
        launch(i % nbooks, streams[i % len(streams)].handle)
    t1 = time.perf_counter()
    for s, ev in zip(streams[1:], joins):
        ev.record(s.handle)
        streams[0].wait(ev)
    e1.record(streams[0].handle)
    e1.sync()
    capi.call("bs_device_sync")
    ms = e0.elapsed_ms(e1) / steps
    if tag:
        LAUNCH_CPU_US[tag] = (t1 - t0) / steps * 1e6
    for ev in [e0, e1] + joins:
        ev.destroy()
    return ms


def reference_cpu(book, threads: int, port_prices) -> dict:
    """Times the reference itself: oracle/_ref/blackscholes_ff, compiled from
    the reference's blackscholes.c with -DENABLE_FF (the FastFlow ParallelFor
    map, NUM_RUNS = 100 passes, BLOCKING_MODE), on `threads` host threads.
    Built without the PARSEC hooks it prints no ROI time, so the ROI is timed
    from outside: its stdout is a pty (so line-buffered), the ROI begins right
    after it prints "Size of data:" (blackscholes.c:769-782, including the
    map's construction as the reference's own ROI does) and ends when it
    creates the output file (:913-925).  The output is then checked against
    the port's prices.  Returns {} when the binary is absent."""
    import pty
    import select
    import shutil
    import tempfile
    import pyoracle
    exe = os.path.join(ROOT, "oracle", "_ref", "blackscholes_ff")
    if not os.access(exe, os.X_OK):
        return {}
    d = tempfile.mkdtemp(prefix="bs_ref_")
    try:
        inp, outp = os.path.join(d, "in.txt"), os.path.join(d, "out.txt")
        pyoracle.write_parsec_input(inp, book)
        master, slave = pty.openpty()
        p = subprocess.Popen([exe, str(threads), inp, outp], stdin=subprocess.DEVNULL,
                             stdout=slave, stderr=slave, close_fds=True)
        os.close(slave)
        text, t_begin, t_end = b"", None, None
        deadline = time.perf_counter() + 600
        while time.perf_counter() < deadline:
            if t_begin is None:
                r, _, _ = select.select([master], [], [], 0.05)
                if r:
                    try:
                        chunk = os.read(master, 4096)
                    except OSError:            # the child closed the pty
                        chunk = b""
                    text += chunk
                    if b"Size of data" in text and text.rstrip(b" ").endswith(b"\n"):
                        t_begin = time.perf_counter()
                if p.poll() is not None and t_begin is None:
                    break
            else:
                if os.path.exists(outp):
                    t_end = time.perf_counter()
                    break
                if p.poll() is not None:
                    break
                time.sleep(0.001)
        rc = p.wait(timeout=600)
        while True:                            # drain what is left on the pty
            r, _, _ = select.select([master], [], [], 0)
            if not r:
                break
            try:
                chunk = os.read(master, 4096)
            except OSError:
                break
            if not chunk:
                break
            text += chunk
        os.close(master)
        if rc != 0 or t_begin is None or t_end is None:
            raise RuntimeError(f"reference run failed (rc {rc}): {text[-400:]!r}")
        with open(outp) as f:
            n_out = int(f.readline())
            got = np.array(f.read().split(), dtype=np.float64).astype(np.float32)
        mism = int(np.count_nonzero(got.view(np.uint32) != port_prices.view(np.uint32)))
        sec = t_end - t_begin
        return {"value": book.n * 100 / sec, "unit": "options/s", "cores": threads,
                "kind": "reference",
                "sample": f"{book.n} options x 100 passes (NUM_RUNS), oracle/_ref/blackscholes_ff "
                          f"(blackscholes.c with -DENABLE_FF) on {threads} threads, ROI "
                          f"{sec:.3f} s timed from its output",
                "outputs_equal_to_port": n_out == book.n and mism == 0}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def cpu_baseline(book_host, passes: int, ref_options: int) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    threads = max(1, min(aff, 16))     # the box's CPU share for one GPU
    _, sec, used = pyoracle.price_f32(book_host, nthreads=threads, passes=passes)
    rate = book_host.n * passes / sec
    one = book_host.slice(0, 1_000_000)
    _, sec1, _ = pyoracle.price_f32(one, nthreads=1, passes=2)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    port = {"value": rate, "unit": "options/s", "cores": used, "kind": "port",
            "sample": f"{book_host.n} options x {passes} passes, static blocks, "
                      f"{used} pinned threads, oracle/bs_oracle.c (bit-exact with "
                      f"blackscholes.c), {sec:.2f} s",
            "single_thread_options_per_s": one.n * 2 / sec1,
            "speedup": rate / (one.n * 2 / sec1),
            "cpu_model": model, "nproc": os.cpu_count(), "affinity": aff}
    ref = {}
    if ref_options > 0:
        sample = book_host.slice(0, min(ref_options, book_host.n))
        port_prices, _, _ = pyoracle.price_f32(sample, nthreads=threads, passes=1)
        try:
            ref = reference_cpu(sample, threads, port_prices)
        except (RuntimeError, OSError, ValueError, subprocess.SubprocessError) as e:
            port["reference_run_failed"] = str(e)[:300]   # the port stands in
            ref = {}
    if not ref:
        return port
    ref.update({"cpu_model": model, "nproc": os.cpu_count(), "affinity": aff, "port": port})
    return ref


def serve_bench(result: dict) -> dict:
    """Runs the serving load generator (p3arsec_amd/bin/bs_serve_bench, a C++
    client of the C ABI) as a child process: open-loop Poisson at 20K req/s
    with direct launches (the default) and with graph replay, open-loop Poisson
    at 1.9M req/s (about three quarters of batch-mode throughput), and
    closed-loop saturation runs (4096 requests outstanding) in both launch
    forms; compares the
    saturated options/s with the batch kernels' options/s for the same 70/30
    price/Greeks mix."""
    exe = os.path.join(ROOT, "p3arsec_amd", "bin", "bs_serve_bench")
    secs = str(ARGS.serve_seconds)
    out = {}
    for tag, extra in (("poisson_20k", ["--rate", "20000"]),
                       ("poisson_20k_graph", ["--rate", "20000", "--use-graph", "1"]),
                       ("poisson_1_9m", ["--rate", "1900000"]),
                       ("saturation", ["--saturate", "--inflight", "4096"]),
                       ("saturation_graph", ["--saturate", "--inflight", "4096", "--use-graph", "1"])):
        p = subprocess.run([exe, "--duration-s", secs, "--budget-gb", "32"] + extra,
                           capture_output=True, text=True, timeout=300)
        if p.returncode != 0:
            raise RuntimeError(f"bs_serve_bench {tag} failed: {p.stderr[-2000:]}")
        out[tag] = json.loads(p.stdout.strip().splitlines()[-1])
    price_rate = result["value"]
    greeks_rate = result.get("greeks", {}).get("options_per_s")
    if greeks_rate:
        batch = 1.0 / (0.7 / price_rate + 0.3 / greeks_rate)
        out["batch_mode_options_per_s_70_30"] = batch
        out["saturation_frac_of_batch"] = out["saturation"]["options_per_s"] / batch
        out["saturation_graph_frac_of_batch"] = out["saturation_graph"]["options_per_s"] / batch
        out["poisson_1_9m_frac_of_batch"] = out["poisson_1_9m"]["options_per_s"] / batch
    return out


SW_TRIALS = 1_000_000    # PARSEC's native run: 128 swaptions, 1,000,000 trials each
SW_COUNT = 128


def swaptions_bench(no_cpu: bool, calls: int = 5) -> dict:
    """The swaptions map on PARSEC's native input: every path of every
    swaption on the GPU (bs_swaptions_host_f64), timed per whole call (host
    records in, prices out; the records built once before, as the reference
    builds its parm array before its ROI) and by HIP events on its stream; parity of the
    whole portfolio at 250,000 trials against the oracle, whose static-block
    run over the host's threads is the CPU baseline."""
    from p3arsec_amd import swaptions as W
    sws, base = W.parsec_portfolio(SW_COUNT)
    paths = SW_COUNT * ((SW_TRIALS + W.BLOCK_SIZE - 1) // W.BLOCK_SIZE * W.BLOCK_SIZE)
    W.price(sws[:4], 10_000, base)            # first call: code object load
    prep = W.Prepared(sws)                    # the records, built once as the driver does
    st, e0, e1 = capi.Stream(), capi.Event(), capi.Event()
    for _ in range(6):                        # clocks settle over ~20 ms of load
        W.price(prep, SW_TRIALS, base, stream=st.handle)
    wall, gpu = [], []
    for _ in range(calls):
        t0 = time.perf_counter()
        e0.record(st.handle)
        p, se = W.price(prep, SW_TRIALS, base, stream=st.handle)
        e1.record(st.handle)
        e1.sync()
        wall.append(time.perf_counter() - t0)
        gpu.append(e0.elapsed_ms(e1))
    out = {"workload": f"{SW_COUNT} swaptions x {SW_TRIALS} trials (PARSEC native input): "
                       "HJM, 11 steps, 3 factors, FP64",
           "seconds_per_call": min(wall), "gpu_ms_per_call": min(gpu),
           "median_seconds_per_call": float(np.median(wall)), "timed_calls": calls,
           "paths_per_s": paths / min(wall), "swaptions_per_s": SW_COUNT / min(wall),
           "price_min": float(p.min()), "price_max": float(p.max())}
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            prof = json.load(f)["swaptions"]
    except (OSError, KeyError, ValueError):
        prof = {}
    if prof.get("paths") == paths:
        # the paths kernel is bound by FP64 VALU issue: its committed
        # instruction count's issue time against the whole call's HIP-event time
        floor_ms = prof["valu_issue_us"] / 1e3
        out["valu"] = {"issue_floor_ms": floor_ms, "frac_of_call": floor_ms / min(gpu),
                       "source": "profiles/hbm_traffic.json (rocprofv3 --pmc SQ_INSTS_VALU)"}
    st.destroy()
    e0.destroy()
    e1.destroy()
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyswaptions
        trials = 250_000
        gp, ge = W.price(sws, trials, base)
        cp, ce, sec, used = pyswaptions.price_batch(sws, base, trials, W.BLOCK_SIZE, 0)
        rel = np.abs(gp - cp) / np.maximum(np.abs(cp), 1e-300)
        out["parity_vs_oracle"] = {"trials": trials,
                                   "max_rel_price_diff": float(rel[cp > 0].max()),
                                   "max_abs_price_diff": float(np.abs(gp - cp).max()),
                                   "max_rel_stderr_diff": float((np.abs(ge - ce) /
                                                                 np.maximum(ce, 1e-300))[ce > 0].max())}
        cpaths = SW_COUNT * trials
        out["cpu_baseline"] = {"value": cpaths / sec, "unit": "paths/s", "cores": used,
                               "kind": "port",
                               "sample": f"{SW_COUNT} swaptions x {trials} trials, static blocks "
                                         f"over swaptions, {used} pinned threads, "
                                         f"oracle/swaptions_oracle.c (HJM_Swaption_Blocking "
                                         f"restated, every time step simulated), {sec:.2f} s",
                               "gpu_speedup": (paths / min(wall)) / (cpaths / sec)}
    return out


def only_one(args, L, host, dbooks, outs, streams) -> int:
    """Runs one kernel for --warmup + --steps passes (profiler runs)."""
    n, nb = host.n, len(dbooks)
    extra = []
    if args.only == "copy":
        # calibration for the HBM traffic counters: device-to-device copies of
        # exactly 256 MiB (read 256 MiB, write 256 MiB per copy)
        nbytes = 256 << 20
        src, dst = capi.DeviceBuffer(nbytes), capi.DeviceBuffer(nbytes)
        capi.call("bs_memset", src.ptr, 1, nbytes, None)
        fn = lambda i, s: L.bs_memcpy_d2d(dst.ptr, src.ptr, nbytes, s)  # noqa
        ms = timed_passes(lambda i, s: capi.check(fn(i, s), "copy"), 1, args.steps, args.warmup,
                          streams[:1])
        print(json.dumps({"kernel": "copy", "bytes": nbytes, "ms_per_step": ms,
                          "gbs_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return 0
    if args.only == "price":
        fn = lambda i, s: L.bs_price_f32(*dbooks[i].ptrs(), outs[i].ptr, n, s)  # noqa
    elif args.only == "fast":
        fn = lambda i, s: L.bs_price_fast_f32(*dbooks[i].ptrs(), outs[i].ptr, n, s)  # noqa
    elif args.only in ("portfolio", "portfolio_greeks"):
        g = 1 if args.only == "portfolio_greeks" else 0
        extra = [capi.DeviceBuffer(40) for _ in range(nb)]
        fn = lambda i, s: L.bs_portfolio_f32(*dbooks[i].ptrs(), None, n, g,  # noqa
                                             extra[i].ptr, s)
    elif args.only == "greeks":
        extra = [[capi.alloc_out(n) for _ in range(5)] for _ in range(nb)]
        fn = lambda i, s: L.bs_price_greeks_f32(*dbooks[i].ptrs(),  # noqa
                                                *[o.ptr for o in extra[i]], n, s)
    else:
        h64 = host.astype(np.float64)
        d64 = [capi.DeviceBook(h64, dtype=np.float64) for _ in range(2)]
        o64 = [capi.alloc_out(n, np.float64) for _ in range(2)]
        extra = [d64, o64]
        nb = 2
        fn = lambda i, s: L.bs_price_f64(*d64[i].ptrs(), o64[i].ptr, n, s)  # noqa
    ss = streams[:1] if args.one_stream else streams[:STREAMS.get(args.only, 2)]
    ms = timed_passes(lambda i, s: capi.check(fn(i, s), args.only), nb, args.steps, args.warmup,
                      ss)
    print(json.dumps({"kernel": args.only, "options": n, "ms_per_step": ms}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--books", type=int, default=4)
    ap.add_argument("--cpu-passes", type=int, default=100)
    ap.add_argument("--cpu-ref-options", type=int, default=4_000_000,
                    help="options in the compiled reference's timed sample (0: port only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--no-swaptions", action="store_true")
    ap.add_argument("--only", choices=["price", "fast", "greeks", "fp64", "copy", "portfolio",
                                       "portfolio_greeks", "swaptions"],
                    help="time one kernel only (for profiler runs)")
    ap.add_argument("--one-stream", action="store_true",
                    help="with --only: every pass on one stream")
    args = ap.parse_args()
    global ARGS
    ARGS = args
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the map needs no collectives: each GPU would price a contiguous range "
              "of the book on its own)",
              file=sys.stderr)
        return 2

    capi.load()
    if capi.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        return 1
    capi.call("bs_set_device", 0)
    if args.only == "swaptions":
        r = swaptions_bench(True, calls=max(1, args.steps))
        print(json.dumps({"kernel": "swaptions", **r}), flush=True)
        return 0
    n, nb = args.n, args.books
    log(f"bench: generating {n} options")
    host = books.make_book(n, books.SEED_BASE + 2)
    dbooks = [capi.DeviceBook(host) for _ in range(nb)]
    outs = [capi.alloc_out(n) for _ in range(nb)]
    streams = [capi.Stream(), capi.Stream()]
    L = capi.load()

    def launch_price(i, s):
        capi.check(L.bs_price_f32(*dbooks[i].ptrs(), outs[i].ptr, n, s), "price")

    spin_up()
    if args.only:
        return only_one(args, L, host, dbooks, outs, streams)
    log("bench: price map")
    ms = timed_passes(launch_price, nb, args.steps, args.warmup, streams[:STREAMS["price"]],
                      "price")
    # the same passes on one stream (each pass waits for the previous one)
    ms_one = timed_passes(launch_price, nb, args.steps, min(args.warmup, 200), streams[:1])
    value = n / (ms * 1e-3)
    alg_bytes = 25 * n
    achieved = alg_bytes / (ms * 1e-3) / 1e9

    # correctness of what was timed (sampled against the oracle)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    got = outs[0].download(np.float32, n)
    idx = np.random.default_rng(0).choice(n, 200_000, replace=False)
    sub = host.slice(0, n)
    sub = books.Book(*(a[idx] for a in sub.arrays()))
    ref, _, _ = pyoracle.price_f32(sub)
    max_err = float(np.max(np.abs(got[idx].astype(np.float64) - ref)))

    result = {
        "metric": METRIC, "value": value, "unit": "options/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "Black-Scholes price map, 10M-option FP32 book, SoA "
                               "(5 x f32 + u8 in, f32 out), one launch per pass",
                   "options": n, "rotating_books": nb, "bytes_per_option": 25,
                   "streams": STREAMS["price"], "max_abs_err_vs_oracle_sample": max_err},
        "one_stream": {"ms_per_step": ms_one, "options_per_s": n / (ms_one * 1e-3),
                       "roofline_frac": alg_bytes / (ms_one * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("price", n)},
    }
    prof = profile_record("price", n)
    if "valu_issue_us" in prof:
        # VALU busy time per pass from the committed SQ_INSTS_VALU count (wave64
        # instructions x 4 cycles over 1024 SIMDs at 2.4 GHz), against this pass time
        result["valu"] = {"issue_us_per_pass": prof["valu_issue_us"],
                          "frac_of_pass": prof["valu_issue_us"] / (ms * 1e3),
                          "source": "profiles/hbm_traffic.json (rocprofv3 --pmc SQ_INSTS_VALU)"}

    if not args.no_extra:
        try:
            # the whole timed book against the reference (the bit-exact FP32
            # port) and against two FP64 oracles: the reference's algorithm in
            # double, and the exact closed form
            log("bench: accuracy of the 10M book")
            ref_full, _, _ = pyoracle.price_f32(host)
            g64 = got.astype(np.float64)
            acc = {}
            for name, want in (("reference_fp32", ref_full),
                               ("fp64_restatement", pyoracle.price_f64(host.astype(np.float64))),
                               ("fp64_closed_form", pyoracle.price_exact(host))):
                d = np.abs(g64 - want)
                acc[name] = {"max_abs_err": float(d.max()), "mean_abs_err": float(d.mean())}
            acc["bit_identical_to_reference"] = float(np.mean(got == ref_full))
            result["accuracy_10m"] = acc

            def launch_fast(i, s):
                capi.check(L.bs_price_fast_f32(*dbooks[i].ptrs(), outs[i].ptr, n, s), "fast")
            log("bench: fast price map")
            fast_ms = timed_passes(launch_fast, nb, args.steps, min(args.warmup, 200),
                                   streams[:STREAMS["fast"]])
            result["fast_variant"] = {"ms_per_step": fast_ms, "options_per_s": n / (fast_ms * 1e-3),
                                      "achieved_gbs": alg_bytes / (fast_ms * 1e-3) / 1e9}
            gouts = [[capi.alloc_out(n) for _ in range(5)] for _ in range(nb)]

            def launch_greeks(i, s):
                capi.check(L.bs_price_greeks_f32(*dbooks[i].ptrs(), *[o.ptr for o in gouts[i]],
                                                 n, s), "greeks")
            log("bench: greeks map")
            spin_up(0.1)
            gms = timed_passes(launch_greeks, nb, max(10, args.steps // 2), min(args.warmup, 200),
                               streams[:STREAMS["greeks"]], "greeks")
            result["greeks"] = {"ms_per_step": gms, "ratio_to_price": gms / ms,
                                "ratio_to_price_one_stream": gms / ms_one,
                                "options_per_s": n / (gms * 1e-3),
                                "achieved_gbs": 41 * n / (gms * 1e-3) / 1e9}
            for row in gouts:
                for o in row:
                    o.free()
            # FP64 map: one book copy in double (2 rotating copies = 980 MB)
            log("bench: fp64 map")
            h64 = host.astype(np.float64)
            d64 = [capi.DeviceBook(h64, dtype=np.float64) for _ in range(2)]
            o64 = [capi.alloc_out(n, np.float64) for _ in range(2)]

            def launch_f64(i, s):
                capi.check(L.bs_price_f64(*d64[i].ptrs(), o64[i].ptr, n, s), "f64")
            spin_up(0.1)   # the uploads above left the GPU idle
            fms = timed_passes(launch_f64, 2, max(10, args.steps // 4), min(args.warmup, 500),
                               streams[:STREAMS["fp64"]])
            result["fp64"] = {"ms_per_step": fms, "ratio_to_fp32": fms / ms,
                              "options_per_s": n / (fms * 1e-3),
                              "achieved_gbs": 49 * n / (fms * 1e-3) / 1e9}
            for b in d64:
                b.free()
            for o in o64:
                o.free()
            # portfolio sums over the same book (the MapReduce step after the map):
            # 21 B/option read, nothing per option written
            log("bench: portfolio sums")
            sums = [capi.DeviceBuffer(40) for _ in range(nb)]

            def launch_pf(i, s, greeks):
                capi.check(L.bs_portfolio_f32(*dbooks[i].ptrs(), None, n, greeks, sums[i].ptr, s),
                           "portfolio")
            tpf = timed_passes(lambda i, s: launch_pf(i, s, 0), nb, max(10, args.steps // 2),
                               min(args.warmup, 200), streams[:STREAMS["price"]])
            tpg = timed_passes(lambda i, s: launch_pf(i, s, 1), nb, max(10, args.steps // 4),
                               min(args.warmup, 100), streams[:STREAMS["price"]])
            result["portfolio"] = {"price_sum_us": tpf * 1e3, "price_sum_gbs": 21 * n / (tpf * 1e-3) / 1e9,
                                   "greeks_sums_us": tpg * 1e3,
                                   "greeks_sums_gbs": 21 * n / (tpg * 1e-3) / 1e9}
            for b in sums:
                b.free()
            # the 4096-option book of the reference's CPU-runnable case: L2-resident, so a
            # per-pass latency, not an HBM figure; error against the FP64 closed form
            log("bench: 4096-option book")
            small = books.make_book(4096, books.SEED_BASE + 1)
            dsm, osm = capi.DeviceBook(small), capi.alloc_out(4096)
            t4 = timed_passes(lambda i, s: capi.check(L.bs_price_f32(*dsm.ptrs(), osm.ptr, 4096, s),
                                                      "4096"), 1, 100, 100, streams[:1])
            # the 100 passes as one call: replayed from a captured graph
            sp = dsm.ptrs()
            t4g = timed_passes(lambda i, s: capi.check(L.bs_price_passes_f32(
                *sp[:5], sp[5], 1, osm.ptr, 4096, 100, s), "4096 graph"), 1, 20, 5, streams[:1]) / 100
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle
            p4 = osm.download(np.float32, 4096).astype(np.float64)
            e4 = np.abs(p4 - pyoracle.price_exact(small))
            r4, _, _ = pyoracle.price_f32(small)
            result["config_4096"] = {"us_per_pass": t4g * 1e3, "options_per_s": 4096 / (t4g * 1e-3),
                                     "passes": 100, "form": "bs_price_passes_f32 (100 passes, graph replay)",
                                     "us_per_pass_launch_by_launch": t4 * 1e3,
                                     "max_err_vs_fp64_closed_form": float(e4.max()),
                                     "mean_err_vs_fp64_closed_form": float(e4.mean()),
                                     "max_abs_diff_vs_reference": float(np.abs(p4 - r4).max())}
            dsm.free()
            osm.free()
            # one GPU's share of an 8-way sharded book (10M/8 options, one book, not
            # rotated): the working set fits the 256 MB MALL, so this is not HBM-bound
            log("bench: 1.25M-option shard")
            ns = n // 8
            shard = capi.DeviceBook(host.slice(0, ns))
            oshard = capi.alloc_out(ns)
            tsh = timed_passes(lambda i, s: capi.check(L.bs_price_f32(*shard.ptrs(), oshard.ptr, ns, s),
                                                       "shard"), 1, 200, 200, streams[:1])
            result["shard_1_25m"] = {"us_per_pass": tsh * 1e3, "options_per_s": ns / (tsh * 1e-3),
                                     "note": "one book reused every pass: MALL-resident working set"}
            shard.free()
            oshard.free()
            # host-buffer boundary (the SkePU2 Map form): upload, one pass, download
            log("bench: host-buffer map")
            arrs = [np.ascontiguousarray(a) for a in host.arrays()[:5]]
            ty32 = host.otype.astype(np.int32)
            hout = np.empty(n, np.float32)
            hts = []
            for _ in range(3):
                t0 = time.perf_counter()
                capi.check(L.bs_price_host_f32(*[a.ctypes.data for a in arrs], ty32.ctypes.data,
                                               hout.ctypes.data, n, 1, 0), "host")
                hts.append(time.perf_counter() - t0)
            result["host_buffers"] = {"seconds_per_call": min(hts),
                                      "options_per_s": n / min(hts),
                                      "note": "pageable host arrays in and out (PCIe-inclusive), "
                                              "one pass; not the headline value"}
        except Exception as e:  # recorded, so the headline line still prints
            section_failed(result, "extras", e)

    if not args.no_extra and not args.no_serve:
        try:
            log("bench: serving loop")
            result["serve"] = serve_bench(result)
        except Exception as e:  # recorded, so the headline line still prints
            section_failed(result, "serve", e)

    if not args.no_extra and not args.no_swaptions:
        try:
            log("bench: swaptions")
            result["swaptions"] = swaptions_bench(args.no_cpu)
        except Exception as e:  # recorded, so the headline line still prints
            section_failed(result, "swaptions", e)

    if not args.no_cpu:
        try:
            log("bench: cpu baseline")
            result["cpu_baseline"] = cpu_baseline(host, args.cpu_passes, args.cpu_ref_options)
        except Exception as e:  # recorded, so the headline line still prints
            section_failed(result, "cpu_baseline", e)

    result["launch_cpu_us_per_step"] = LAUNCH_CPU_US
    print(json.dumps(result), flush=True)
    for b in dbooks:
        b.free()
    for o in outs:
        o.free()
    for s in streams:
        s.destroy()
    return 0


if __name__ == "__main__":
    sys.exit(main())
