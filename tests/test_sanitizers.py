"""CPU: the library's host C++ (HNSW builder with several threads, the
hnswlib / NSG / FPGA-layout file formats, the data generator) built with
AddressSanitizer and UndefinedBehaviorSanitizer, and separately with
ThreadSanitizer (tests/san/Makefile), and run
over round trips, the reference-written golden indexes and the error paths
for malformed files (tests/san/host_sanitize.cpp).  Any sanitizer report
fails the run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None or not os.path.isdir(os.path.join(HERE, "san")),
                    reason="needs g++, make and tests/san")
def test_host_code_under_asan_and_ubsan(tmp_path):
    out = str(tmp_path / "bin")
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "san"), f"OUT={out}"], check=True)
    run = tmp_path / "run"
    run.mkdir()
    args = [os.path.join(out, "host_sanitize"), str(run)]
    for name, dim, metric in (("l2_n1000_d16_M8", 16, 0), ("ip_n1000_d24_M6", 24, 1), ("l2_n400_d8_M4", 8, 0)):
        args += [os.path.join(GOLD, name + ".hnsw.bin"), str(dim), str(metric)]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None or not os.path.isdir(os.path.join(HERE, "san")),
                    reason="needs g++, make and tests/san")
def test_host_code_under_tsan(tmp_path):
    """The multithreaded builder (4 threads) and the rest under ThreadSanitizer."""
    out = str(tmp_path / "bin")
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "san"), f"OUT={out}", "tsan"], check=True)
    run = tmp_path / "run"
    run.mkdir()
    r = subprocess.run([os.path.join(out, "host_tsan"), str(run)], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0 and "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
