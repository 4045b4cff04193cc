"""The serving loop's concurrency on the CPU (no GPU): the dispatcher thread,
the clients' calls and shutdown of p3arsec_amd/csrc/serve_loop.hpp driven by
tests/native/serve_loop_stress.cpp with a host executor in place of the GPU.

Each scenario exits non-zero if an invariant fails: every ticket finishes
exactly once; batch outputs lie in the arena and overlap neither each other,
nor another batch in flight, nor a resident book; injected launch and batch
failures reach exactly the requests of that batch and free their outputs;
shutdown finishes the batches in flight and fails the rest; admission waits
for released results when the arena is full; the arena returns to the books'
bytes when every result is released; with lanes on, every batch holds one
lane's requests in that lane's slots, and small requests finish while a slow
bulk batch is still running; the context's ticket lock excludes and serves
eight threads in turn, one of them sleeping while it holds it.  The same
scenarios run again under ThreadSanitizer."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
SCENARIOS = ["stress", "faults", "shutdown", "books", "memory", "lanes", "lock"]


def build(target):
    subprocess.run(["make", "-s", "-C", NATIVE, target], check=True)
    return os.path.join(NATIVE, target)


def run(exe, scenario, timeout):
    p = subprocess.run([exe, scenario], capture_output=True, text=True, timeout=timeout)
    print(p.stdout.strip())
    assert p.returncode == 0, p.stdout + p.stderr[-4000:]
    return p.stdout


@pytest.mark.parametrize("scenario", SCENARIOS)
def test_loop(scenario):
    out = run(build("serve_loop_stress"), scenario, 120)
    if scenario == "faults":
        assert " 0 failed" not in out
    if scenario == "memory":
        assert " 0 admission waits" not in out


@pytest.mark.skipif(not os.path.exists("/opt/rocm/llvm/bin/clang++") and shutil.which("clang++") is None,
                    reason="no clang for the ThreadSanitizer build")
@pytest.mark.parametrize("scenario", SCENARIOS)
def test_loop_under_thread_sanitizer(scenario):
    exe = build("serve_loop_stress_tsan")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    p = subprocess.run([exe, scenario], capture_output=True, text=True, timeout=300, env=env)
    print(p.stdout.strip())
    assert p.returncode == 0, p.stderr[-6000:]
    assert "ThreadSanitizer" not in p.stderr
