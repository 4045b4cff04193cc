"""Benchmark of the Black-Scholes map on one MI355X.

    python bench.py [--gpus 1] [--steps 500] [--warmup 2000] [--prewarm-ms 250]

A step is one pass of the price map over the 10,000,000-option FP32 book
(SoA: s, k, r, v, t float32 + uint8 type in, float32 price out; 25 bytes per
option), one kernel launch per pass, as the reference's NUM_RUNS loop
(blackscholes.c:316-347).  Inputs are resident in HBM before timing starts.
Passes rotate over --books device copies of the book (4 x 250 MB), so the
256 MB MALL cannot hold the working set between passes, and alternate
between two HIP streams, so the last waves of one pass overlap the first
waves of the next (passes are independent: adjacent passes use different
books); the same passes on one stream are reported beside it.

Every timed section first runs its own kernel on its own rotating buffers for
--prewarm-ms of wall time (recorded under prewarm_ms): a kernel that starts on
an idle or differently loaded GPU runs its first 10-20 ms 15-20% slow, and 20
passes of the 10M book last under 1 ms.  Then --warmup untimed passes, then
exactly --steps passes between HIP events.  The reference's 100 passes are
timed as one more region right after (passes_100).

Prints one JSON line: options/s over the timed passes, the roofline of the
price kernel (algorithmic bytes / average launch time vs 8 TB/s), and the
CPU baseline (the compiled reference's FastFlow build on the host's cores,
and the oracle's static-block port beside it).  Extra sections: accuracy of
the whole 10M book, the fused Greeks kernel (time ratio to price-only and
per-Greek errors at 10M), the FP64 map (time and error distribution), the
portfolio sums, the 4096-option book, the serving loop and swaptions.
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
# on MI355X (10M options, 300 passes): price 37.25 / 37.52 us per pass on two
# streams vs 37.98 / 39.41 on one (two boxes); Greeks 67.22 vs 67.98; FP64
# 85 vs 89.  Over 20 passes the two schedules of the price map are within
# about 1 us of each other, either way round depending on the box.
STREAMS = {"price": 2, "fast": 2, "greeks": 2, "fp64": 2}


def reap_children() -> int:
    """Kills and waits for any child process still alive (every section waits
    for its own children; this is the backstop).  Returns how many there were."""
    try:
        import psutil
    except ImportError:
        return -1
    kids = psutil.Process().children(recursive=True)
    for c in kids:
        try:
            c.kill()
        except psutil.Error:
            pass
    psutil.wait_procs(kids, timeout=10)
    return len(kids)


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


PREWARM_MS = {}


def prewarm(launch, nbooks: int, streams, ms: float, tag: str) -> None:
    """Runs the section's own kernel on its rotating books, launches
    alternating over `streams`, for `ms` of wall time before the section's
    warm-up and timed passes.  A kernel that starts on an idle or differently
    loaded GPU runs its first 10-20 ms through a power-management dip (per-pass
    times 15-20% above the settled ones); a short timed region (20 passes of
    the 10M book are under 1 ms) would otherwise measure the dip.  The time
    spent is recorded under prewarm_ms in the JSON line."""
    t0 = time.perf_counter()
    i = 0
    while (time.perf_counter() - t0) * 1e3 < ms:
        for _ in range(4 * nbooks):
            launch(i % nbooks, streams[i % len(streams)].handle)
            i += 1
        capi.call("bs_device_sync")
    PREWARM_MS[tag] = round((time.perf_counter() - t0) * 1e3, 1)


def timed_passes(launch, nbooks: int, steps: int, warmup: int, streams, tag: str = "") -> float:
    """Runs warmup + steps passes, pass i on book i % nbooks and stream
    i % len(streams) (launch(i, stream)); returns ms per pass of the timed
    passes.  After a host synchronize, the warmup passes are enqueued, every
    stream is joined on the device into the stream of the first timed pass,
    then the start event on that stream, the timed passes (every other stream
    waits for the start event first), a join of every stream into the stream
    of the last timed pass and the end event on it; a host synchronize
    follows.  The start event fires when the last warmup pass ends, with the
    timed launches already queued behind it (the warmup passes cover the
    host's launch time), so the region holds exactly the timed passes and no
    host launch latency; and the first timed pass follows the start event on
    its own stream, so no cross-stream wait sits at either end of the region.
    The events are timing-only (bs_event_create_timing); the host synchronize
    after the region makes the results visible.  Passes on different streams
    touch different books and outputs (nbooks is a multiple of the stream
    count), so with two streams one pass's last waves overlap the next pass's
    first ones.  Also records the host time per launch call
    (LAUNCH_CPU_US[tag])."""
    assert nbooks % len(streams) == 0
    ns = len(streams)
    first, last = streams[warmup % ns], streams[(warmup + steps - 1) % ns]
    capi.call("bs_device_sync")
    for i in range(warmup):
        launch(i % nbooks, streams[i % ns].handle)
    # timing-only events: no system-scope fence (cache writeback and
    # invalidation) at the markers inside or at the ends of the region
    e0, e1 = capi.Event(timing=True), capi.Event(timing=True)
    pre = [capi.Event(timing=True) for _ in streams[1:]]
    joins = [capi.Event(timing=True) for _ in streams[1:]]
    for s, ev in zip([x for x in streams if x is not first], pre):   # warmups end before e0
        ev.record(s.handle)
        first.wait(ev)
    e0.record(first.handle)
    for s in streams:
        if s is not first:
            s.wait(e0)
    t0 = time.perf_counter()
    for i in range(warmup, warmup + steps):
        launch(i % nbooks, streams[i % ns].handle)
    t1 = time.perf_counter()
    for s, ev in zip([x for x in streams if x is not last], joins):
        ev.record(s.handle)
        last.wait(ev)
    e1.record(last.handle)
    e1.sync()
    capi.call("bs_device_sync")
    ms = e0.elapsed_ms(e1) / steps
    if tag:
        LAUNCH_CPU_US[tag] = (t1 - t0) / steps * 1e6
    for ev in [e0, e1] + pre + joins:
        ev.destroy()
    return ms


def reference_roi(exe: str, threads: int, inp: str, outp: str) -> float:
    """One run of the compiled reference on `threads` threads: returns the
    seconds of its region of interest.  Built without the PARSEC hooks (whose
    hooks.c needs a configure-generated config.h and the Mammut library) it
    prints no ROI time, so the ROI is timed from outside: its stdout is a pty
    (so line-buffered), the ROI begins right after it prints "Size of data:"
    (blackscholes.c:769-782, including the map's construction as the
    reference's own ROI does) and ends when it creates the output file
    (:913-925); both are polled every 50 us.  The child is always waited for
    and the pty closed, also when the run fails."""
    import pty
    import select
    if os.path.exists(outp):
        os.unlink(outp)
    master, slave = pty.openpty()
    p = None
    try:
        p = subprocess.Popen([exe, str(threads), inp, outp], stdin=subprocess.DEVNULL,
                             stdout=slave, stderr=slave, close_fds=True)
        os.close(slave)
        slave = -1
        text, t_begin, t_end = b"", None, None
        deadline = time.perf_counter() + 600
        while time.perf_counter() < deadline:
            if t_begin is None:
                r, _, _ = select.select([master], [], [], 0.00005)
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
                time.sleep(0.00005)
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
    finally:
        if p is not None and p.poll() is None:
            p.kill()
            p.wait()
        if slave >= 0:
            os.close(slave)
        os.close(master)
    if rc != 0 or t_begin is None or t_end is None:
        raise RuntimeError(f"reference run failed (rc {rc}): {text[-400:]!r}")
    return t_end - t_begin


def read_reference_output(outp: str, n: int):
    with open(outp) as f:
        n_out = int(f.readline())
        got = np.array(f.read().split(), dtype=np.float64).astype(np.float32)
    return got if n_out == n and got.size == n else None


def reference_cpu(book, threads: int, port_prices, repeats: int = 1) -> dict:
    """Times the reference itself: oracle/_ref/blackscholes_ff, compiled from
    the reference's blackscholes.c with -DENABLE_FF (the FastFlow ParallelFor
    map, NUM_RUNS = 100 passes, BLOCKING_MODE), on `threads` host threads,
    `repeats` times (the best and median ROI are reported).  The output is
    checked against the port's prices.  Returns {} when the binary is absent."""
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
        secs = [reference_roi(exe, threads, inp, outp) for _ in range(repeats)]
        got = read_reference_output(outp, book.n)
        same = got is not None and bool(np.array_equal(got.view(np.uint32),
                                                       port_prices.view(np.uint32)))
        sec = min(secs)
        return {"value": book.n * 100 / sec, "unit": "options/s", "cores": threads,
                "kind": "reference",
                "sample": f"{book.n} options x 100 passes (NUM_RUNS), oracle/_ref/blackscholes_ff "
                          f"(blackscholes.c with -DENABLE_FF) on {threads} threads, ROI "
                          f"{sec:.4f} s timed from its output"
                          + (f" (best of {repeats}; median {float(np.median(secs)):.4f} s)"
                             if repeats > 1 else ""),
                "roi_seconds": secs, "outputs_equal_to_port": same}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def cpu_model() -> str:
    """lscpu's model name (the full string) with its CPU family, model and
    stepping numbers (a virtual machine may report only "AMD EPYC"; the numbers
    still name the generation), else /proc/cpuinfo's model name."""
    try:
        p = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10)
        f = {}
        for ln in p.stdout.splitlines():
            k, _, v = ln.partition(":")
            f.setdefault(k.strip(), v.strip())
        if f.get("Model name"):
            ids = ", ".join(f"{k.lower()} {f[k]}" for k in ("CPU family", "Model", "Stepping")
                            if f.get(k))
            return f"{f['Model name']} ({ids})" if ids else f["Model name"]
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            return next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        return ""


def cpu_baseline(book_host, passes: int, ref_options: int) -> dict:
    """The reference's CPU path on every logical core this process may run on
    (the GPU box's share of its host): the compiled reference (FastFlow build)
    on the 10M book and on the 4096-option book of the reference's own
    CPU-runnable case, and the oracle's static-block port beside it (1 thread
    and all threads)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    threads = max(1, aff)
    _, sec, used = pyoracle.price_f32(book_host, nthreads=threads, passes=passes)
    rate = book_host.n * passes / sec
    one = book_host.slice(0, 1_000_000)
    _, sec1, _ = pyoracle.price_f32(one, nthreads=1, passes=2)
    model = cpu_model()
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
            small = books.make_book(4096, books.SEED_BASE + 1)
            sp, _, _ = pyoracle.price_f32(small)
            c1 = reference_cpu(small, threads, sp, repeats=7)
            c1_one = reference_cpu(small, 1, sp, repeats=3)
            if c1:
                ref["config_4096"] = {"ms_100_passes": min(c1["roi_seconds"]) * 1e3,
                                      "ms_100_passes_median": float(np.median(c1["roi_seconds"])) * 1e3,
                                      "options_per_s": c1["value"], "threads": threads,
                                      "ms_100_passes_1_thread": min(c1_one["roi_seconds"]) * 1e3,
                                      "outputs_equal_to_port": c1["outputs_equal_to_port"],
                                      "note": "4096 options x 100 passes, the ROI including the "
                                              "FastFlow map's construction (blackscholes.c:782-913)"}
        except (RuntimeError, OSError, ValueError, subprocess.SubprocessError) as e:
            port["reference_run_failed"] = str(e)[:300]   # the port stands in
            ref = {}
    if not ref:
        return port
    ref.update({"cpu_model": model, "nproc": os.cpu_count(), "affinity": aff, "port": port})
    return ref


def serve_bench(result: dict) -> dict:
    """Runs the serving load generator (p3arsec_amd/bin/bs_serve_bench, a C++
    client of the C ABI) as a child process, once per load, each on a catalog
    of 1024 resident books log-uniform over [1, 2^20] options (80.9M
    options, 1.7 GB of inputs: far more than the 256 MB MALL, so the loop
    streams its books from HBM as the batch benchmark does): open-loop Poisson
    at 20K req/s with direct launches (the default form) and with graph
    replay, open-loop Poisson at 2.0M req/s (about 73% of the batch kernels'
    options/s: the mean request is 75.6K options), and closed-loop saturation (4096
    requests outstanding) in both launch forms.  The saturated options/s is
    compared with the batch kernels' options/s for the same 70/30
    price/Greeks mix (their rotating 10M books, also HBM-resident)."""
    exe = os.path.join(ROOT, "p3arsec_amd", "bin", "bs_serve_bench")
    secs = str(ARGS.serve_seconds)
    out = {}
    for tag, extra in (("poisson_20k", ["--rate", "20000"]),
                       ("poisson_20k_graph", ["--rate", "20000", "--use-graph", "1"]),
                       ("poisson_2m", ["--rate", "2000000"]),
                       ("saturation", ["--saturate", "--inflight", "4096"]),
                       ("saturation_graph", ["--saturate", "--inflight", "4096", "--use-graph", "1"])):
        p = subprocess.run([exe, "--duration-s", secs, "--budget-gb", "32", "--books", "1024"]
                           + extra, capture_output=True, text=True, timeout=300)
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
        out["poisson_2m_frac_of_batch"] = out["poisson_2m"]["options_per_s"] / batch
    d, g = out["poisson_20k"], out["poisson_20k_graph"]
    out["launch_form"] = {
        "default": "direct launch (bs_serve_opts.use_graph = 0)",
        "p99_ms_20k": {"direct": d["p99_ms"], "graph_replay": g["p99_ms"]},
        "p99_ms_le64k_20k": {"direct": d["p99_ms_le64k"], "graph_replay": g["p99_ms_le64k"]},
        "saturation_options_per_s": {"direct": out["saturation"]["options_per_s"],
                                     "graph_replay": out["saturation_graph"]["options_per_s"]},
        "why": "graph replay costs more host time per launch than the three direct launches "
               "it replaces, which sets the low-load latency; at saturation both forms reach "
               "the same throughput"}
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
        # the same restatement by chains: only the time steps the payoff reads
        # (the pruning the GPU kernel does), the same bits
        _, _, sec_c, _ = pyswaptions.price_batch(sws, base, trials, W.BLOCK_SIZE, 0, chain=True)
        gpu_rate = paths / min(wall)
        out["cpu_baseline"] = {"value": cpaths / sec, "unit": "paths/s", "cores": used,
                               "kind": "port",
                               "sample": f"{SW_COUNT} swaptions x {trials} trials, static blocks "
                                         f"over swaptions, {used} pinned threads, "
                                         f"oracle/swaptions_oracle.c (HJM_Swaption_Blocking "
                                         f"restated, every time step simulated), {sec:.2f} s",
                               "pruned_port_paths_per_s": cpaths / sec_c,
                               "pruned_port_note": "the same restatement per forward-rate chain, "
                                                   "simulating only the time steps the payoff "
                                                   f"reads, as the GPU kernel does: {sec_c:.2f} s",
                               "gpu_speedup_vs_unpruned_port": gpu_rate / (cpaths / sec),
                               "gpu_speedup_vs_pruned_port": gpu_rate / (cpaths / sec_c)}
    return out


GREEK_NAMES = ("price", "delta", "vega", "rho", "theta")


def accuracy_10m(host, got) -> dict:
    """The whole timed book against the reference (the bit-exact FP32 port of
    blackscholes.c:190-258) and two FP64 oracles: the reference's algorithm in
    double, and the exact closed form.  over_1e-4 counts the prices the
    reference's ERR_CHK test (fabs(delta) >= 1e-4, blackscholes.c:333-340)
    would flag against the reference's own output."""
    import pyoracle
    ref_full, _, _ = pyoracle.price_f32(host)
    g64 = got.astype(np.float64)
    r64 = ref_full.astype(np.float64)
    f64 = pyoracle.price_f64(host.astype(np.float64))
    exact = pyoracle.price_exact(host)
    acc = {}
    for name, want in (("reference_fp32", r64), ("fp64_restatement", f64),
                       ("fp64_closed_form", exact)):
        d = np.abs(g64 - want)
        acc[name] = {"max_abs_err": float(d.max()), "mean_abs_err": float(d.mean())}
    acc["over_1e-4_vs_reference"] = int(np.count_nonzero(np.abs(g64 - r64) >= 1e-4))
    acc["bit_identical_to_reference"] = float(np.mean(got == ref_full))
    acc["reference_itself"] = {
        "vs_fp64_restatement_max": float(np.abs(r64 - f64).max()),
        "vs_fp64_closed_form_max": float(np.abs(r64 - exact).max()),
        "vs_fp64_closed_form_mean": float(np.abs(r64 - exact).mean())}
    return acc


def fast_section(L, dbooks, outs, n, streams) -> dict:
    """Fast32: hardware transcendentals (not within 1e-4 of the reference on
    every option; reported for the cost of the reference's roundings)."""
    nb = len(dbooks)

    def launch(i, s):
        capi.check(L.bs_price_fast_f32(*dbooks[i].ptrs(), outs[i].ptr, n, s), "fast")
    ss = streams[:STREAMS["fast"]]
    prewarm(launch, nb, ss, ARGS.prewarm_ms / 2, "fast")
    ms = timed_passes(launch, nb, ARGS.steps, ARGS.warmup, ss)
    return {"ms_per_step": ms, "options_per_s": n / (ms * 1e-3),
            "achieved_gbs": 25 * n / (ms * 1e-3) / 1e9}


def greek_errors(g, book, chk) -> dict:
    """Per-Greek errors of g (5 x n: price, delta, vega, rho, theta) against
    chk: max relative |g - chk| / max(|chk|, 1e-3 scale) and max/mean scaled
    |g - chk| / scale, scales S+K, 1, S sqrt(T), K T, S+K."""
    s, k, t = (a.astype(np.float64) for a in (book.s, book.k, book.t))
    scales = (s + k, np.ones_like(s), s * np.sqrt(t), k * t, s + k)
    out = {}
    for q, name in enumerate(GREEK_NAMES):
        d = np.abs(g[q] - chk[q])
        rel = d / np.maximum(np.abs(chk[q]), 1e-3 * scales[q])
        sc = d / scales[q]
        out[name] = {"max_rel": float(rel.max()), "max_scaled": float(sc.max()),
                     "mean_scaled": float(sc.mean())}
    return out


def greeks_section(L, host, dbooks, n, streams, ms, ms_one) -> dict:
    """The fused price + Greeks kernel over the rotating 10M books: time per
    pass and its ratio to the price map, then every option's five outputs
    against FP64 central differences of the A&S price (bump 1e-3 relative)
    and the closed-form Greeks."""
    import pyoracle
    nb = len(dbooks)
    gouts = [[capi.alloc_out(n) for _ in range(5)] for _ in range(nb)]

    def launch(i, s):
        capi.check(L.bs_price_greeks_f32(*dbooks[i].ptrs(), *[o.ptr for o in gouts[i]], n, s),
                   "greeks")
    ss = streams[:STREAMS["greeks"]]
    prewarm(launch, nb, ss, ARGS.prewarm_ms, "greeks")
    gms = timed_passes(launch, nb, ARGS.steps, ARGS.warmup, ss, "greeks")
    gms100 = timed_passes(launch, nb, 100, 4, ss)
    g = np.stack([o.download(np.float32, n) for o in gouts[0]])
    for row in gouts:
        for o in row:
            o.free()
    out = {"ms_per_step": gms, "ratio_to_price": gms / ms, "ratio_to_price_one_stream": gms / ms_one,
           "ms_per_pass_100": gms100, "options_per_s": n / (gms * 1e-3),
           "achieved_gbs": 41 * n / (gms * 1e-3) / 1e9,
           "roofline_frac": 41 * n / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    g64 = g.astype(np.float64)
    out["errors_10m"] = {"fd_bump_1e-3": greek_errors(g64, host, pyoracle.greeks_fd(host, 1e-3)),
                         "closed_form": greek_errors(g64, host, pyoracle.greeks_exact(host))}
    return out


def fp64_section(L, host, got, n, streams, ms) -> dict:
    """fptype = double over the same 10M book (two rotating device copies,
    980 MB): time per pass, and the error distribution of the FP64 kernel
    against the FP64 oracles and of the FP32 kernel against the FP64 kernel."""
    import pyoracle
    h64 = host.astype(np.float64)
    d64 = [capi.DeviceBook(h64, dtype=np.float64) for _ in range(2)]
    o64 = [capi.alloc_out(n, np.float64) for _ in range(2)]

    def launch(i, s):
        capi.check(L.bs_price_f64(*d64[i].ptrs(), o64[i].ptr, n, s), "f64")
    ss = streams[:STREAMS["fp64"]]
    prewarm(launch, 2, ss, ARGS.prewarm_ms / 2, "fp64")
    fms = timed_passes(launch, 2, ARGS.steps, ARGS.warmup, ss)
    p64 = o64[0].download(np.float64, n)
    for b in d64:
        b.free()
    for o in o64:
        o.free()
    want = pyoracle.price_f64(h64)
    exact = pyoracle.price_exact(host)
    g32 = got.astype(np.float64)

    def stats(a, b):
        d = np.abs(a - b)
        rel = d / np.maximum(np.abs(b), 1e-300)
        return {"max_abs": float(d.max()), "mean_abs": float(d.mean()),
                "p99_abs": float(np.quantile(d, 0.99)),
                "max_rel_where_price_ge_0.01": float(rel[np.abs(b) >= 0.01].max())}
    return {"ms_per_step": fms, "ratio_to_fp32": fms / ms, "options_per_s": n / (fms * 1e-3),
            "achieved_gbs": 49 * n / (fms * 1e-3) / 1e9,
            "roofline_frac": 49 * n / (fms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "errors_10m": {"fp64_kernel_vs_fp64_restatement": stats(p64, want),
                           "fp64_kernel_vs_fp64_closed_form": stats(p64, exact),
                           "fp32_kernel_vs_fp64_kernel": stats(g32, p64),
                           "fp32_kernel_vs_fp64_closed_form": stats(g32, exact)}}


def portfolio_section(L, dbooks, n, streams, ms) -> dict:
    """Portfolio sums over the same books (the MapReduce step after the map):
    21 B/option read, nothing per option written."""
    nb = len(dbooks)
    sums = [capi.DeviceBuffer(40) for _ in range(nb)]
    out = {}
    for greeks, tag in ((0, "price_sum"), (1, "greeks_sums")):
        def launch(i, s):
            capi.check(L.bs_portfolio_f32(*dbooks[i].ptrs(), None, n, greeks, sums[i].ptr, s),
                       "portfolio")
        ss = streams[:STREAMS["price"]]
        prewarm(launch, nb, ss, ARGS.prewarm_ms / 2, "portfolio_" + tag)
        t = timed_passes(launch, nb, ARGS.steps, ARGS.warmup, ss)
        out[tag + "_us"] = t * 1e3
        out[tag + "_gbs"] = 21 * n / (t * 1e-3) / 1e9
    out["price_map_us"] = ms * 1e3
    for b in sums:
        b.free()
    return out


def config4096_section(L, streams) -> dict:
    """The 4096-option book of the reference's CPU-runnable case, 100 passes:
    L2-resident, so a per-pass latency, not an HBM figure; errors against the
    FP64 closed form and the reference."""
    import pyoracle
    small = books.make_book(4096, books.SEED_BASE + 1)
    dsm, osm = capi.DeviceBook(small), capi.alloc_out(4096)
    t4 = timed_passes(lambda i, s: capi.check(L.bs_price_f32(*dsm.ptrs(), osm.ptr, 4096, s),
                                              "4096"), 1, 100, 100, streams[:1])
    # the 100 passes as one call: replayed from a captured graph
    sp = dsm.ptrs()
    t4g = timed_passes(lambda i, s: capi.check(L.bs_price_passes_f32(
        *sp[:5], sp[5], 1, osm.ptr, 4096, 100, s), "4096 graph"), 1, 20, 5, streams[:1]) / 100
    p4 = osm.download(np.float32, 4096).astype(np.float64)
    e4 = np.abs(p4 - pyoracle.price_exact(small))
    r4, _, _ = pyoracle.price_f32(small)
    dsm.free()
    osm.free()
    return {"us_per_pass": t4g * 1e3, "options_per_s": 4096 / (t4g * 1e-3),
            "ms_100_passes": t4g * 100, "passes": 100,
            "form": "bs_price_passes_f32 (100 passes, graph replay)",
            "us_per_pass_launch_by_launch": t4 * 1e3,
            "max_err_vs_fp64_closed_form": float(e4.max()),
            "mean_err_vs_fp64_closed_form": float(e4.mean()),
            "max_abs_diff_vs_reference": float(np.abs(p4 - r4).max())}


def shard_section(L, host, n, streams) -> dict:
    """One GPU's share of an 8-way sharded book (10M/8 options, one book, not
    rotated): the working set fits the 256 MB MALL, so this is not HBM-bound."""
    ns = n // 8
    shard = capi.DeviceBook(host.slice(0, ns))
    oshard = capi.alloc_out(ns)
    tsh = timed_passes(lambda i, s: capi.check(L.bs_price_f32(*shard.ptrs(), oshard.ptr, ns, s),
                                               "shard"), 1, 200, 200, streams[:1])
    shard.free()
    oshard.free()
    return {"us_per_pass": tsh * 1e3, "options_per_s": ns / (tsh * 1e-3),
            "note": "one book reused every pass: MALL-resident working set"}


def repack_section(L, host, n, streams) -> dict:
    """The reference's AoS -> SoA repack (BS:747-767) on the device over the
    10M book: 36-byte OptionData records in, five float arrays and the type
    bytes out (57 B per option), two rotating record buffers (720 MB) on two
    streams, timed like the map; the output of the last pass is checked
    against the book bit for bit."""
    rec = np.dtype([("s", "<f4"), ("strike", "<f4"), ("r", "<f4"), ("divq", "<f4"),
                    ("v", "<f4"), ("t", "<f4"), ("OptionType", "S1"), ("pad", "V3"),
                    ("divs", "<f4"), ("DGrefval", "<f4")])
    aos = np.zeros(n, rec)
    for f, a in (("s", host.s), ("strike", host.k), ("r", host.r), ("v", host.v),
                 ("t", host.t)):
        aos[f] = a
    aos["OptionType"] = np.where(host.otype != 0, b"P", b"C")
    daos = [capi.DeviceBuffer.from_array(aos.view(np.uint8)) for _ in range(2)]
    del aos
    outs = [[capi.DeviceBuffer(n * 4) for _ in range(5)] + [capi.DeviceBuffer(n)]
            for _ in range(2)]

    def launch(i, st):
        capi.check(L.bs_repack_aos_f32(daos[i].ptr, n, *[o.ptr for o in outs[i]], st), "repack")
    ss = streams[:2]
    prewarm(launch, 2, ss, ARGS.prewarm_ms / 2, "repack_aos")
    t = timed_passes(launch, 2, ARGS.steps, ARGS.warmup, ss)
    last = (ARGS.warmup + ARGS.steps - 1) % 2
    same = all(np.array_equal(o.download(np.float32, n).view(np.uint32), a.view(np.uint32))
               for o, a in zip(outs[last][:5], (host.s, host.k, host.r, host.v, host.t)))
    same = same and np.array_equal(outs[last][5].download(np.uint8, n),
                                   (host.otype != 0).astype(np.uint8))
    for b in daos + [o for row in outs for o in row]:
        b.free()
    gbs = 57 * n / (t * 1e-3) / 1e9
    return {"us_per_pass": t * 1e3, "options_per_s": n / (t * 1e-3), "achieved_gbs": gbs,
            "roofline_frac": gbs / HBM_PEAK_GBS, "bytes_per_option": 57,
            "output_equals_book": bool(same)}


def host_buffers_section(L, host, n) -> dict:
    """The host-buffer boundary (the SkePU2 Map form): upload, one pass,
    download."""
    arrs = [np.ascontiguousarray(a) for a in host.arrays()[:5]]
    ty32 = host.otype.astype(np.int32)
    hout = np.empty(n, np.float32)
    hts = []
    for _ in range(3):
        t0 = time.perf_counter()
        capi.check(L.bs_price_host_f32(*[a.ctypes.data for a in arrs], ty32.ctypes.data,
                                       hout.ctypes.data, n, 1, 0), "host")
        hts.append(time.perf_counter() - t0)
    return {"seconds_per_call": min(hts), "options_per_s": n / min(hts),
            "note": "pageable host arrays in and out (PCIe-inclusive), one pass; not the "
                    "headline value"}


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
    if args.only == "repack":
        r = repack_section(L, host, n, streams)
        print(json.dumps({"kernel": "repack", "options": n, **r,
                          "prewarm_ms": PREWARM_MS.get("repack_aos")}), flush=True)
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
    launch = lambda i, s: capi.check(fn(i, s), args.only)  # noqa
    prewarm(launch, nb, ss, args.prewarm_ms, args.only)
    ms = timed_passes(launch, nb, args.steps, args.warmup, ss)
    print(json.dumps({"kernel": args.only, "options": n, "ms_per_step": ms,
                      "prewarm_ms": PREWARM_MS.get(args.only)}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--prewarm-ms", type=float, default=250.0,
                    help="wall time each timed section first runs its own kernel for")
    ap.add_argument("--books", type=int, default=4)
    ap.add_argument("--cpu-passes", type=int, default=100)
    ap.add_argument("--cpu-ref-options", type=int, default=10_000_000,
                    help="options in the compiled reference's timed sample (0: port only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--no-swaptions", action="store_true")
    ap.add_argument("--only", choices=["price", "fast", "greeks", "fp64", "copy", "portfolio",
                                       "portfolio_greeks", "swaptions", "repack"],
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
    s_price = streams[:STREAMS["price"]]
    prewarm(launch_price, nb, s_price, args.prewarm_ms, "price")
    ms = timed_passes(launch_price, nb, args.steps, args.warmup, s_price, "price")
    # the reference's NUM_RUNS = 100 passes, timed as one region right after
    ms100 = timed_passes(launch_price, nb, 100, 4, s_price)
    # the same passes on one stream (each pass waits for the previous one: the
    # region / passes is the kernel's average launch duration on its stream)
    prewarm(launch_price, nb, streams[:1], 50, "price_one_stream")
    ms_one = timed_passes(launch_price, nb, args.steps, args.warmup, streams[:1])
    ms_one100 = timed_passes(launch_price, nb, 100, 4, streams[:1])
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
                   "streams": len(s_price), "max_abs_err_vs_oracle_sample": max_err},
        "passes_100": {"ms_per_pass": ms100, "options_per_s": n / (ms100 * 1e-3),
                       "roofline_frac": alg_bytes / (ms100 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "note": "the reference's NUM_RUNS = 100 passes as one timed region, "
                               "right after the --steps region"},
        "one_stream": {"ms_per_step": ms_one, "options_per_s": n / (ms_one * 1e-3),
                       "roofline_frac": alg_bytes / (ms_one * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "ms_per_pass_100": ms_one100,
                       "roofline_frac_100": alg_bytes / (ms_one100 * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("price", n),
                     "kernel": "k_stream<PriceF32<Ref32, 2>> (bs_price_f32)",
                     "duration_us": ms * 1e3,
                     "one_stream_launch_us": ms_one * 1e3,
                     "streams": len(s_price),
                     "note": "duration_us is the time per launch with launches alternating over "
                             "two streams (each overlaps its neighbours, so a trace's start-to-end "
                             "average per dispatch spans about two passes and the trace's window "
                             "per dispatch is the comparable figure); one_stream_launch_us is the "
                             "launch's own average duration with every launch on one stream"},
    }
    prof = profile_record("price", n)
    if "valu_issue_us" in prof:
        # VALU busy time per pass from the committed SQ_INSTS_VALU count (wave64
        # instructions x 4 cycles over 1024 SIMDs at 2.4 GHz), against this pass time
        result["valu"] = {"issue_us_per_pass": prof["valu_issue_us"],
                          "lane_instructions_per_option": prof.get("valu_per_option"),
                          "frac_of_pass": prof["valu_issue_us"] / (ms * 1e3),
                          "source": "profiles/hbm_traffic.json (rocprofv3 --pmc SQ_INSTS_VALU)"}

    if not args.no_extra:
        for name, fn in (("accuracy_10m", lambda: accuracy_10m(host, got)),
                         ("fast_variant", lambda: fast_section(L, dbooks, outs, n, streams)),
                         ("greeks", lambda: greeks_section(L, host, dbooks, n, streams, ms, ms_one)),
                         ("fp64", lambda: fp64_section(L, host, got, n, streams, ms)),
                         ("portfolio", lambda: portfolio_section(L, dbooks, n, streams, ms)),
                         ("config_4096", lambda: config4096_section(L, streams)),
                         ("shard_1_25m", lambda: shard_section(L, host, n, streams)),
                         ("repack_aos_10m", lambda: repack_section(L, host, n, streams)),
                         ("host_buffers", lambda: host_buffers_section(L, host, n))):
            try:
                log(f"bench: {name}")
                result[name] = fn()
            except Exception as e:  # recorded, so the headline line still prints
                section_failed(result, name, e)

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
    result["prewarm_ms"] = PREWARM_MS
    result["children_left"] = reap_children()
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
