"""The product's host code on the CPU under sanitizers: the topology code
(grid.cpp, topo_api.cpp) under AddressSanitizer, LeakSanitizer and
UndefinedBehaviorSanitizer (long seeded refine/compress sequences with slot
reuse, table builds, Morton ordering and ghost expansion,
tests/host/sanitize_topology.cpp), and the host worker pool under
ThreadSanitizer.  Any out-of-bounds access, use after free, leak, data race
or undefined operation ends the driver with a sanitizer report."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gliomauq_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_topology_host_code_under_sanitizers():
    d = tempfile.mkdtemp(prefix="gqsan_")
    try:
        exe = os.path.join(d, "sanitize_topology")
        cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
               "-fno-sanitize-recover=undefined", os.path.join(ROOT, "tests", "host", "sanitize_topology.cpp"),
               os.path.join(CSRC, "grid.cpp"), os.path.join(CSRC, "topo_api.cpp"), "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-4000:]
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
                   UBSAN_OPTIONS="print_stacktrace=1")
        r = subprocess.run([exe, "60"], capture_output=True, text=True, timeout=600, env=env)
        assert r.returncode == 0 and r.stdout.strip() == "ok", (r.stdout[-2000:], r.stderr[-6000:])
    finally:
        shutil.rmtree(d, ignore_errors=True)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_worker_pool_under_thread_sanitizer():
    """The host worker pool (gliomauq_amd/csrc/pool.hpp) under ThreadSanitizer
    (tests/host/pool_stress.cpp): pools destroyed right after construction
    must shut down (a worker that first ran after the destructor's wake-up
    used to sleep for ever and the join never returned), every loop item is
    covered exactly once, and a chunk's exception reaches the caller."""
    d = tempfile.mkdtemp(prefix="gqpool_")
    try:
        exe = os.path.join(d, "pool_stress")
        cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", "-fsanitize=thread",
               os.path.join(ROOT, "tests", "host", "pool_stress.cpp"), "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-4000:]
        r = subprocess.run([exe, "2000"], capture_output=True, text=True, timeout=120,
                           env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
        assert r.returncode == 0 and r.stdout.strip() == "ok", (r.stdout[-2000:], r.stderr[-6000:])
    finally:
        shutil.rmtree(d, ignore_errors=True)
