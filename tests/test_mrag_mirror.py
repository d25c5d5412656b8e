"""The C++ mirror of the reference's block-operator API (include/gq_mrag.hpp)
takes a driver written with the reference's own calls: tests/host/
mrag_driver.cpp follows Glioma_HG_UQ.h:16-50, Glioma_HG_UQ.cpp:266-277 and
:511-588 (Grid<W,B>, BlockCollection, BoundaryInfo, pipeline_process +
static process, Science::AutomaticRefinement<0,0>(grid, blockfwt, tol,
maxLevel, 1, &profiler), AutomaticCompression<0,0>).  It compiles on the CPU;
on a GPU its adaptive run is bit-identical to the Python mirror's run of the
same schedule."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host", "mrag_driver.cpp")
INC = os.path.join(ROOT, "include")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_reference_shaped_driver_compiles():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", INC, SRC],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_reference_shaped_driver_runs_like_python_mirror():
    import gliomauq_amd as gq
    from gliomauq_amd import anatomy as A
    d = tempfile.mkdtemp(prefix="gqdrv_")
    try:
        exe = os.path.join(d, "mrag_driver")
        libdir = os.path.join(ROOT, "gliomauq_amd")
        r = subprocess.run(["g++", "-std=c++17", "-O2", "-I", INC, SRC, "-o", exe, "-L", libdir, "-lgq",
                            "-Wl,-rpath," + libdir], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        n, L = 64, 6.4
        pwm, pgm, u0, _ = A.config_inputs(n, L, 3)
        u0 = (u0 * 5).astype(np.float32)
        bpd, maxl, tolr, tolc, dumpfreq, tend = n // 16, 2, 1e-3, 1e-4, 2.0, 12.0
        ctx = gq.Context(0)
        g = gq.Grid(ctx, bpd, maxl, 4)
        g.sample_volume("phi", u0)
        g.sample_volume("pw", pwm)
        g.sample_volume("pg", pgm)
        s = gq.Glioma_HG_UQ(g, L, cfl=0.99, adaptive=True, tol_refine=tolr, tol_compress=tolc, dumpfreq=dumpfreq)
        Dw, Dg, rho, dt = s.params(0.02, 0.08)
        steps, _ = s.run(0.02, 0.08, tend)
        want = g.dump_uniform()
        blocks = g.n_blocks()
        ctx.close()
        out = os.path.join(d, "u.bin")
        stdin = b"".join(np.ascontiguousarray(v, np.float32).tobytes() for v in (u0, pwm, pgm))
        r = subprocess.run([exe, str(bpd), str(maxl), "4", "%.9g" % Dw, "%.9g" % rho, repr(tend), repr(dumpfreq),
                            repr(tolr), repr(tolc), out], input=stdin, capture_output=True, timeout=300)
        assert r.returncode == 0, r.stderr.decode()
        got = np.fromfile(out, np.float64).reshape(n, n, n)
        print("driver: %s; python mirror: steps %d blocks %d" % (r.stdout.decode().strip(), steps, blocks))
        assert r.stdout.decode().split() == ["steps", str(steps), "blocks", str(blocks)]
        assert got.tobytes() == want.tobytes()
        assert blocks != bpd ** 3  # the run adapted the grid
    finally:
        shutil.rmtree(d, ignore_errors=True)
