"""Bindings of the reference-from-source artefacts in oracle/_ref (TEST ONLY).

oracle/_ref holds the reference's own code compiled from /root/reference by
oracle/Makefile (`make -C oracle ref`): the FK operator (Real=float and
Real=double), the MRAG ghost-table builder, the refine/compress planners and
the likelihood program.
Tests skip when these are absent.
"""
import ctypes as C
import os
import shutil
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")


def available(name):
    return os.path.exists(os.path.join(REF_DIR, name))


def _load(name):
    return C.CDLL(os.path.join(REF_DIR, name))


def fk_block(lab_phi, lab_pw, lab_pg, h, Dw, Dg, rho, real="f64"):
    L = _load("libgqref_fk_%s.so" % real)
    dt = np.float64 if real == "f64" else np.float32
    ct = C.c_double if real == "f64" else C.c_float
    ptr = np.ctypeslib.ndpointer(dt, flags="C_CONTIGUOUS")
    f = getattr(L, "gqref_fk_block_%s" % real)
    f.argtypes = [ptr, ptr, ptr, C.c_int, ct, ct, ct, ct, ptr]
    B = lab_phi.shape[0] - 2
    out = np.zeros((B, B, B), dt)
    rc = f(*(np.ascontiguousarray(a, dt) for a in (lab_phi, lab_pw, lab_pg)), B, h, Dw, Dg, rho, out)
    assert rc == 0
    return out


def update(phi, dphidt, dt_step, real="f64"):
    L = _load("libgqref_fk_%s.so" % real)
    dt = np.float64 if real == "f64" else np.float32
    ptr = np.ctypeslib.ndpointer(dt, flags="C_CONTIGUOUS")
    f = getattr(L, "gqref_update_%s" % real)
    f.argtypes = [ptr, ptr, C.c_int, C.c_double]
    p = np.ascontiguousarray(phi, dt).copy()
    rc = f(p, np.ascontiguousarray(dphidt, dt), phi.shape[0], dt_step)
    assert rc == 0
    return p


def uniform_steps(phi, pw, pg, h, Dw, Dg, rho, dt, n_steps, threads=0, real="f32"):
    """n_steps of the reference's RD functor + UpdateTumor (compiled from
    source) on a uniform periodic N^3 grid of 16^3 blocks, cells with the
    CCell_extended footprint, OpenMP over blocks; volumes [z, y, x]."""
    L = _load("libgqref_fk_%s.so" % real)
    dt_ = np.float64 if real == "f64" else np.float32
    ct = C.c_double if real == "f64" else C.c_float
    ptr = np.ctypeslib.ndpointer(dt_, flags="C_CONTIGUOUS")
    f = getattr(L, "gqref_uniform_steps_%s" % real)
    f.argtypes = [ptr, ptr, ptr, C.c_int, C.c_double, ct, ct, ct, C.c_double, C.c_int, C.c_int]
    p = np.ascontiguousarray(phi, dt_).copy()
    rc = f(p, np.ascontiguousarray(pw, dt_), np.ascontiguousarray(pg, dt_), p.shape[0], h, Dw, Dg, rho, dt, n_steps,
           threads)
    if rc != 0:
        raise RuntimeError("reference uniform driver failed: %d" % rc)
    return p


def ghost_terms(levels, indices, target, order, B, s0, s1, cap=1 << 22):
    """Reference MRAG_BBInfoCreator instruction lists for one block's lab.

    Returns (offsets, leaf, cell, weight) over the non-interior lab points in
    z,y,x order."""
    L = _load("libgqref_ghosts.so")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.gqref_ghost_terms.argtypes = [C.c_int, ip, ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    ip, ip, ip, dp]
    lv = np.ascontiguousarray(levels, np.int32)
    ix = np.ascontiguousarray(indices, np.int32).reshape(-1)
    Lw = B + s1 - s0 - 1
    ng = Lw ** 3 - B ** 3
    off = np.zeros(ng + 1, np.int32)
    leaf = np.zeros(cap, np.int32)
    cell = np.zeros(cap, np.int32)
    w = np.zeros(cap, np.float64)
    n = L.gqref_ghost_terms(len(lv), lv, ix, target, order, B, s0, s1, cap, off, leaf, cell, w)
    if n < 0:
        raise RuntimeError("reference ghost harness failed: %d" % n)
    assert n == ng
    m = off[ng]
    return off, leaf[:m].copy(), cell[:m].copy(), w[:m].copy()


def _plan_lib():
    L = _load("libgqref_plan.so")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    u8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
    L.gqref_plan_refine.argtypes = [C.c_int, ip, ip, u8, C.c_int, C.c_int, ip]
    L.gqref_plan_compress.argtypes = [C.c_int, ip, ip, u8, C.c_int, ip]
    return L


def plan_refine(keys, flagged, max_level):
    """Leaves (level, x, y, z) that the reference's Refiner_SpaceExtension(1,
    max_level) plan splits, for the leaf list `keys` with the set `flagged`."""
    lv = np.ascontiguousarray([k[0] for k in keys], np.int32)
    ix = np.ascontiguousarray([k[1:] for k in keys], np.int32).reshape(-1)
    fl = np.ascontiguousarray([k in flagged for k in keys], np.uint8)
    out = np.zeros(len(keys) + 1, np.int32)
    n = _plan_lib().gqref_plan_refine(len(keys), lv, ix, fl, max_level, len(out), out)
    if n < 0:
        raise RuntimeError("reference refine planner harness failed: %d" % n)
    return {keys[i] for i in out[:n]}


def plan_compress(keys, flagged):
    """Parents (level, x, y, z) that the reference's Compressor(1) plan
    collapses, for the leaf list `keys` with the set `flagged`."""
    lv = np.ascontiguousarray([k[0] for k in keys], np.int32)
    ix = np.ascontiguousarray([k[1:] for k in keys], np.int32).reshape(-1)
    fl = np.ascontiguousarray([k in flagged for k in keys], np.uint8)
    out = np.zeros(4 * (len(keys) + 1), np.int32)
    n = _plan_lib().gqref_plan_compress(len(keys), lv, ix, fl, len(keys) + 1, out)
    if n < 0:
        raise RuntimeError("reference compress planner harness failed: %d" % n)
    return {tuple(int(v) for v in out[4 * i:4 * i + 4]) for i in range(n)}


def write_matrix3d(path, arr):
    """Matrix::D3D<double>::dump (Glioma/Matrix.h:609-637): header {1234,3,nx,ny,nz}, typeId 0, x-fastest."""
    nz, ny, nx = arr.shape
    with open(path, "wb") as f:
        np.array([1234, 3, nx, ny, nz, 0], np.int32).tofile(f)
        np.ascontiguousarray(arr, np.float64).tofile(f)


def write_matrix2d(path, arr_xy):
    """Matrix::D2D<double>: header {1234,2,nx,ny}, typeId 0, x-fastest (arr_xy[y][x])."""
    ny, nx = arr_xy.shape
    with open(path, "wb") as f:
        np.array([1234, 2, nx, ny, 0], np.int32).tofile(f)
        np.ascontiguousarray(arr_xy, np.float64).tofile(f)


def run_likelihood(model, pet, t1, t2, params, points_xyz=None, selective=1, step_pet=3, step_ti=2):
    """Run the reference likelihood program in a temp dir; returns -log L parsed from Likelihood.txt
    and the stdout (which prints L_Pet, L_T1, L_T2 with %Lf)."""
    exe = os.path.join(REF_DIR, "likelihood")
    d = tempfile.mkdtemp(prefix="gqlik_")
    try:
        write_matrix3d(os.path.join(d, "HGG_data.dat"), model)
        write_matrix3d(os.path.join(d, "PET_data.dat"), pet)
        write_matrix3d(os.path.join(d, "T1_data.dat"), t1)
        write_matrix3d(os.path.join(d, "T2_data.dat"), t2)
        if points_xyz is not None:
            write_matrix2d(os.path.join(d, "Sphere.dat"), points_xyz)
        with open(os.path.join(d, "LikelihoodInput.txt"), "w") as f:
            for p in params:
                f.write("%.17g\n" % p)
        r = subprocess.run([exe, "-selectivePoints", str(selective), "-stepPet", str(step_pet), "-stepTi",
                            str(step_ti)], cwd=d, capture_output=True, text=True, timeout=120)
        if r.returncode != 0:
            raise RuntimeError(r.stdout + r.stderr)
        with open(os.path.join(d, "Likelihood.txt")) as f:
            val = float(f.read().split()[0])
        return val, r.stdout
    finally:
        shutil.rmtree(d, ignore_errors=True)
