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

import gliomauq_amd as gq
from gliomauq_amd import anatomy as A
from oracle import oracle as O
from oracle import patient_oracle as PO
from oracle import topology as T
from oracle.ref import write_matrix2d, write_matrix3d

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gliomauq_amd", "bin")
FIX = os.path.join(ROOT, "tests", "golden", "reference_fixtures.npz")


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


def run_brain(d, tend, Dw=0.02, rho=0.05, ic=(0.45, 0.5, 0.52), flags=("-bpd", "8", "-maxlevel", "3", "-wavelets",
                                                                        "4", "-adaptive", "0"), params_text=None,
              ic_raw=None, verbose=False):
    with open(os.path.join(d, "HGG_InputParameters.txt"), "w") as f:
        f.write(params_text if params_text is not None else "%r %r %r\n" % (Dw, rho, tend))
    raw = ic_raw if ic_raw is not None else np.array(list(ic) + [0.0] * 9, np.float32)
    raw.tofile(os.path.join(d, "HGG_TumorIC.bin"))
    cmd = [os.path.join(BIN, "gq_brain")] + list(flags) + ["-anatomyDir", d + "/", "-pID", "1"]
    if verbose:
        cmd += ["-verbose", "1"]
    r = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    u = read_matrix3d(os.path.join(d, "HGG_data.dat"))
    return (u, r.stdout) if verbose else u


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


def reference_params(Dw_cm, rho, L, n, cfl=0.99):
    """(Dw, Dg, rho, dt) of Glioma_HG_UQ::run (:519-539) through the Python mirror."""
    from types import SimpleNamespace
    return gq.Glioma_HG_UQ(SimpleNamespace(bpd=n // 16), L, cfl=cfl).params(Dw_cm, rho)


def test_gq_brain_adaptive_dumpfreq_against_oracle_schedule():
    """gq_brain -adaptive 1 -dumpfreq D with the reference's default wavelets
    (Interp2) and cfl (0.99) against the FP64 oracle running the reference's
    time schedule (adapt whenever t >= whenToWrite, Glioma_HG_UQ.cpp:541-575)
    with its own decisions: the same step count, the same leaf count after
    every adaptation, and the dumped field within 1e-5.  The tolerances
    (1e-2 / 1e-3) let this 3.2 cm domain compress."""
    d = tempfile.mkdtemp(prefix="gqbrain_")
    gm, wm, csf, t1, t2, _ = anatomy_files(d)
    ic = (0.45, 0.5, 0.52)
    n = 64
    tend, dumpfreq, Dw_cm, rho_in = 2.0, 0.25, 0.02, 0.05
    tolr, tolc = 1e-2, 1e-3
    u, out = run_brain(d, tend, Dw=Dw_cm, rho=rho_in, ic=ic, verbose=True,
                       flags=("-bpd", "4", "-maxlevel", "2", "-adaptive", "1", "-dumpfreq", repr(dumpfreq),
                              "-tolref", repr(tolr), "-tolcomp", repr(tolc)))
    lines = out.splitlines()
    prm = [l for l in lines if l.startswith("params ")][0]
    vals = dict(kv.split("=") for kv in prm.split()[1:])
    Dw, Dg, rho, dt = reference_params(Dw_cm, rho_in, 3.2, n)
    assert float(vals["dt"]) == dt and float(vals["Dw"]) == np.float32(Dw) and float(vals["Dg"]) == np.float32(Dg)
    blocks_cli = [int(l.split("blocks ")[1].rstrip(")")) for l in lines if l.startswith("Dumping data")]
    steps_cli, final_cli = [int(v.strip(",")) for v in lines[-1].split()[1:4:2]]
    phi0, pw, pg = PO.patient_ic(gm, wm, csf, t1, t2, ic, n)
    topo = T.Topology(n // 16, 2, B=16, order=2)
    fld = T.AdaptiveField(topo)
    for name, v in (("phi", phi0), ("pw", pw), ("pg", pg)):
        fld.set_from_volume(name, lambda l, org, h, v=v: v[int(round(org[2] * n)):int(round(org[2] * n)) + 16,
                                                          int(round(org[1] * n)):int(round(org[1] * n)) + 16,
                                                          int(round(org[0] * n)):int(round(org[0] * n)) + 16])
    steps, hist = T.run_adaptive_days(fld, Dw, Dg, rho, dt, tend, dumpfreq, tolr, tolc, 2,
                                      threads=min(16, os.cpu_count() or 1))
    keys, g = topo.leaf_grid()
    ref = g.dump_uniform(fld.stacked("phi", keys).reshape(len(keys), -1), n)
    e2 = float(np.sqrt(np.sum((u - ref) ** 2) / np.sum(ref ** 2)))
    print("gq_brain adaptive: %d steps, leaf counts %s (oracle %s), final %d (oracle %d), dump rel-L2 %.3e"
          % (steps_cli, blocks_cli, hist[:-1], final_cli, hist[-1], e2))
    assert steps_cli == steps and len(blocks_cli) == len(hist) - 1 >= 4
    assert blocks_cli == hist[:-1] and final_cli == hist[-1]
    assert min(hist) < (n // 16) ** 3  # the grid was compressed
    assert e2 <= 1e-5


def synthetic_patient(d, n=128, seed=0):
    """A 128^3-voxel anatomy (1 mm voxels: L = 12.8 cm) of the synthetic
    generator, written as Patient01/P01_{GM,WM,CSF,T1,T2}.dat."""
    pwm, pgm, tissue = A.anatomy(n, n * 0.1, seed)
    vols = {"GM": pgm, "WM": pwm, "CSF": np.zeros_like(pwm), "T1": np.zeros_like(pwm), "T2": np.zeros_like(pwm)}
    pdir = os.path.join(d, "Patient01")
    os.makedirs(pdir)
    xyz = {}
    for name, V in vols.items():
        Vx = np.ascontiguousarray(np.transpose(V.astype(np.float64), (2, 1, 0)))  # [z,y,x] -> [x,y,z]
        write_xyz(os.path.join(pdir, "P01_%s.dat" % name), Vx)
        xyz[name] = Vx
    return xyz


@pytest.mark.skipif(not os.path.exists(FIX), reason="reference fixtures missing")
def test_gq_brain_reference_example_inputs():
    """makefile/Example's HGG_InputParameters.txt (Dw 0.00239 cm^2/day, rho
    0.0175/day, T 597.6 days) and HGG_TumorIC.bin (0.4872, 0.6523, 0.6548) as
    the reference's runHGG.sh hands them to `brain`, on a synthetic 128^3-voxel
    anatomy: gq_brain against the FP64 oracle of the same solve."""
    fx = np.load(FIX)
    d = tempfile.mkdtemp(prefix="gqbrain_")
    V = synthetic_patient(d)
    text = str(fx["input_parameters_text"])
    u = run_brain(d, None, params_text=text, ic_raw=fx["tumor_ic"],
                  flags=("-bpd", "8", "-maxlevel", "3", "-adaptive", "0"))
    n = 128
    phi0, pw, pg = PO.patient_ic(V["GM"], V["WM"], V["CSF"], V["T1"], V["T2"], fx["tumor_ic"][:3], n)
    assert (phi0 > 0).sum() > 5
    Dw_cm, rho_in, tend = fx["input_parameters"]
    Dw, Dg, rho, dt = reference_params(Dw_cm, rho_in, 12.8, n)
    steps = gq.n_steps_until(tend, dt)
    ref = O.fk_uniform_steps(phi0.astype(np.float64), pw.astype(np.float64), pg.astype(np.float64), 16, 1.0 / n,
                             Dw, Dg, rho, dt, steps, threads=min(16, os.cpu_count() or 1))
    e2 = float(np.sqrt(np.sum((u - ref) ** 2) / np.sum(ref ** 2)))
    einf = float(np.max(np.abs(u - ref)))
    print("gq_brain on the reference's example inputs: %d steps, rel-L2 %.3e, Linf %.3e, tumour volume %d voxels"
          % (steps, e2, einf, int((ref > 0.25).sum())))
    assert steps > 500 and (ref > 0.25).sum() > 100
    assert e2 <= 1e-5 and einf <= 1e-5


@pytest.mark.skipif(not os.path.exists(FIX), reason="reference fixtures missing")
def test_gq_likelihood_reference_sphere_and_input():
    """Likelihood/makefile's Sphere.dat (80,160 points) and
    LikelihoodInput.txt, written back as the reference ships them, read by
    gq_likelihood on seeded 128^3 volumes: against the FP64 oracle."""
    fx = np.load(FIX)
    d = tempfile.mkdtemp(prefix="gqlik_")
    rng = np.random.default_rng(3)
    n = 128
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    r = np.sqrt((x - 66.0) ** 2 + (y - 70.0) ** 2 + (z - 60.0) ** 2)
    model = 1.0 / (1.0 + np.exp((r - 25.0) / 4.0))
    pet = np.where(rng.random((n, n, n)) < 0.4, np.clip(model + 0.05 * rng.standard_normal((n, n, n)), 0, None), 0.)
    t1 = ((model > 0.7) ^ (rng.random((n, n, n)) < 0.05)).astype(np.float64)
    t2 = ((model > 0.25) ^ (rng.random((n, n, n)) < 0.05)).astype(np.float64)
    pts = fx["sphere"].astype(np.float64)
    write_matrix3d(os.path.join(d, "HGG_data.dat"), model)
    write_matrix3d(os.path.join(d, "PET_data.dat"), pet)
    write_matrix3d(os.path.join(d, "T1_data.dat"), t1)
    write_matrix3d(os.path.join(d, "T2_data.dat"), t2)
    write_matrix2d(os.path.join(d, "Sphere.dat"), pts)
    head = np.fromfile(os.path.join(d, "Sphere.dat"), np.int32, 5)
    assert head.tolist() == fx["sphere_header"].tolist()
    with open(os.path.join(d, "LikelihoodInput.txt"), "w") as f:
        f.write("\n".join(repr(float(p)) for p in fx["likelihood_input"]) + "\n")
    for sel in (1, 0):
        r = subprocess.run([os.path.join(BIN, "gq_likelihood"), "-selectivePoints", str(sel), "-stepPet", "3",
                            "-stepTi", "2"], cwd=d, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        val = float(open(os.path.join(d, "Likelihood.txt")).read().split()[0])
        ref_l, _ = O.lik_ref(model, pet, t1, t2, fx["likelihood_input"], selective=bool(sel), points=pts)
        print("Sphere.dat selective=%d: gq_likelihood %.6f oracle %.6f" % (sel, val, -ref_l))
        assert abs(val - (-ref_l)) <= 1e-6 + 1e-12 * abs(ref_l)
