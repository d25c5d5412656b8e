"""The drop-in programs of the reference's file contract: gq_brain (for
`brain -anatomy HGG_UQ`, Glioma/Glioma_HG_UQ.cpp) and gq_likelihood (for
`likelihood`, Likelihood/source/Likelihood_main.cpp), run as processes on
files in a temporary directory.

gq_brain is checked against the restated patient-case initialisation
(oracle/patient_oracle.py, _ic_PatientCase) -- bit-exact initial field --
and, after stepping, against the FP64 oracle solve of the same fields
(relative L2 <= 1e-5).  gq_likelihood is checked against the oracle of
HGG_Likelihood (FP64; the file holds "%Lf", 6 decimals).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from oracle import patient_oracle as PO
from oracle.ref import write_matrix2d, write_matrix3d

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gliomauq_amd", "bin")


def read_matrix3d(path):
    h = np.fromfile(path, np.int32, 6)
    assert h[0] == 1234 and h[1] == 3 and h[5] == 0
    nx, ny, nz = h[2:5]
    return np.fromfile(path, np.float64, offset=24).reshape(nz, ny, nx)


def write_xyz(path, V):
    """V indexed [x, y, z] -> Matrix D3D (x fastest)"""
    write_matrix3d(path, np.ascontiguousarray(np.transpose(V, (2, 1, 0))))


def anatomy_files(d, pid=1, seed=0):
    rng = np.random.default_rng(seed)
    sx, sy, sz = 30, 32, 26
    g = np.random.default_rng(seed + 1)
    wm = g.random((sx, sy, sz))
    gm = 1 - wm
    csf = np.zeros((sx, sy, sz))
    csf[12:16, 10:14, 8:12] = 0.8
    brain = np.zeros((sx, sy, sz), bool)
    brain[3:-3, 3:-3, 3:-3] = True
    wm, gm = wm * brain, gm * brain
    t1 = np.zeros((sx, sy, sz))
    t1[14:17, 15:18, 12:15] = 1.0
    t2 = t1.copy()
    pdir = os.path.join(d, "Patient%02d" % pid)
    os.makedirs(pdir)
    for name, V in (("GM", gm), ("WM", wm), ("CSF", csf), ("T1", t1), ("T2", t2)):
        write_xyz(os.path.join(pdir, "P%02d_%s.dat" % (pid, name)), V)
    return gm, wm, csf, t1, t2, rng


def run_brain(d, tend, Dw=0.02, rho=0.05, ic=(0.45, 0.5, 0.52)):
    with open(os.path.join(d, "HGG_InputParameters.txt"), "w") as f:
        f.write("%r %r %r\n" % (Dw, rho, tend))
    np.array(list(ic) + [0.0] * 9, np.float32).tofile(os.path.join(d, "HGG_TumorIC.bin"))
    r = subprocess.run([os.path.join(BIN, "gq_brain"), "-bpd", "8", "-maxlevel", "3", "-wavelets", "4",
                        "-adaptive", "0", "-anatomyDir", d + "/", "-pID", "1"], cwd=d, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return read_matrix3d(os.path.join(d, "HGG_data.dat"))


def test_gq_brain_initial_condition_and_solve():
    d = tempfile.mkdtemp(prefix="gqbrain_")
    gm, wm, csf, t1, t2, _ = anatomy_files(d)
    ic = (0.45, 0.5, 0.52)
    n = 128
    phi0, pw, pg = PO.patient_ic(gm, wm, csf, t1, t2, ic, n)
    assert (phi0 > 0).sum() > 10 and (pw > 0).sum() > 10000
    u0 = run_brain(d, -1.0)  # while (t <= tend) never runs: the dump is the initial field
    assert np.array_equal(u0, phi0.astype(np.float64))
    tend = 0.4
    u = run_brain(d, tend)
    L = 32 * 0.1
    Dw = float(np.float32(np.float32(0.02) / (L * L)))
    Dg = float(np.float32(0.1 * Dw))
    h = 1.0 / n
    dt = 0.99 * h * h / (6 * max(Dw, Dg))
    steps = 0
    t = 0.0
    while t <= tend:
        steps += 1
        t += dt
    ref = O.fk_uniform_steps(phi0.astype(np.float64), pw.astype(np.float64), pg.astype(np.float64), 16, h, Dw, Dg,
                             float(np.float32(0.05)), dt, steps)
    e = float(np.sqrt(np.mean((u - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    print("gq_brain %d steps: rel L2 %.3e" % (steps, e))
    assert steps > 3 and e <= 1e-5


def test_gq_brain_reports_missing_files():
    d = tempfile.mkdtemp(prefix="gqbrain_")
    r = subprocess.run([os.path.join(BIN, "gq_brain")], cwd=d, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "HGG_InputParameters.txt" in r.stderr


def test_gq_likelihood_matches_oracle():
    d = tempfile.mkdtemp(prefix="gqlik_")
    rng = np.random.default_rng(3)
    n = 32
    model = rng.random((n, n, n))
    pet = np.where(rng.random((n, n, n)) < 0.4, rng.random((n, n, n)), 0.0)
    t1 = (rng.random((n, n, n)) < 0.3).astype(np.float64)
    t2 = (rng.random((n, n, n)) < 0.5).astype(np.float64)
    pts = np.unique(rng.integers(0, n // 2, (300, 3)) * 2, axis=0).astype(np.float64)
    params = [0.000361, 0.85, 0.7, 0.25, 0.09]
    write_matrix3d(os.path.join(d, "HGG_data.dat"), model)
    write_matrix3d(os.path.join(d, "PET_data.dat"), pet)
    write_matrix3d(os.path.join(d, "T1_data.dat"), t1)
    write_matrix3d(os.path.join(d, "T2_data.dat"), t2)
    write_matrix2d(os.path.join(d, "Sphere.dat"), np.ascontiguousarray(pts.T))  # Points(i, c): npoints x 3
    with open(os.path.join(d, "LikelihoodInput.txt"), "w") as f:
        f.write(" ".join(repr(p) for p in params) + "\n")
    for sel in (1, 0):
        r = subprocess.run([os.path.join(BIN, "gq_likelihood"), "-selectivePoints", str(sel)], cwd=d,
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        val = float(open(os.path.join(d, "Likelihood.txt")).read().split()[0])
        ref_l, _ = O.lik_ref(model, pet, t1, t2, params, selective=bool(sel), points=np.ascontiguousarray(pts.T))
        print("selective=%d: gq_likelihood %.6f oracle %.6f" % (sel, val, -ref_l))
        assert abs(val - (-ref_l)) <= 1e-6 + 1e-12 * abs(ref_l)
