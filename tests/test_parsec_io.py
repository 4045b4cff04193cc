"""The CLI's PARSEC text I/O (p3arsec_amd/csrc/tools/parsec_io.hpp) on the CPU:
the threaded reader and writer against the reference's fscanf/fprintf calls
(blackscholes.c:701-734, :923-947) on regular, fixed-point, scientific,
CRLF/tab-separated, spread-over-lines, two-per-line, over- and under-counted,
malformed, empty, hex-counted and special-value inputs, with 1-8 threads
(tests/native/parsec_io_check.cpp)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")


def test_threaded_io_matches_fscanf_and_fprintf():
    subprocess.run(["make", "-s", "-C", NATIVE, "parsec_io_check"], check=True)
    p = subprocess.run([os.path.join(NATIVE, "parsec_io_check")], capture_output=True, text=True,
                       timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
