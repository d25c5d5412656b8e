"""Restatement of Glioma_HG_UQ::_ic_PatientCase (Glioma/Glioma_HG_UQ.cpp:67-225)
-- TEST INFRASTRUCTURE ONLY (checks the gq_brain CLI).

Anatomy volumes (GM, WM, CSF, T1, T2 indexed [x, y, z] as MatrixD3D(i, j, k))
mapped onto the vertex points of a uniform grid of gpd^3 points on the unit
cube with the reference's Real = float arithmetic: nearest-lower voxel with
the aspect-ratio shift (:134-142), CSF removed under MRI signal (:161-162),
normalisation and CSF > 0.1 -> pure CSF (:178-194), and the smoothed-ball
tumour of radius 0.005 over two points of h = 1/128 (:198-208).
Returns phi, pw, pg as float32 [z, y, x].
"""
import numpy as np


def patient_ic(GM, WM, CSF, T1, T2, tumor_ic, gpd):
    sx, sy, sz = GM.shape
    smax = max(sx, sy, sz)
    bh = 1.0 / smax
    f32 = np.float32
    h = f32(1.0) / f32(gpd)
    i = np.arange(gpd)
    x1 = (i.astype(f32) * h).astype(np.float64)  # origin + h*i in float, exact for power-of-two spacing
    z, y, x = np.meshgrid(x1, x1, x1, indexing="ij")
    mx = np.floor(x / bh).astype(np.int64) - int((smax - sx) * 0.5)
    my = np.floor(y / bh).astype(np.int64) - int((smax - sy) * 0.5)
    mz = np.floor(z / bh).astype(np.int64) - int((smax - sz) * 0.5)
    inside = (mx >= 0) & (mx < sx) & (my >= 0) & (my < sy) & (mz >= 0) & (mz < sz)
    cx, cy, cz = np.clip(mx, 0, sx - 1), np.clip(my, 0, sy - 1), np.clip(mz, 0, sz - 1)

    def take(V):
        return np.where(inside, V[cx, cy, cz], 0.0).astype(f32)

    PGt, PWt, Pcsf, PT1, PT2 = (take(V) for V in (GM, WM, CSF, T1, T2))
    mri = PT1 + PT2
    Pcsf = np.where(mri > 0, f32(0), Pcsf)
    allv = PWt.astype(np.float64) + PGt.astype(np.float64) + Pcsf.astype(np.float64)
    pos = allv > 0
    safe = np.where(pos, allv, 1.0)
    PGn = np.where(pos, (PGt / safe).astype(f32), PGt)
    PWn = np.where(pos, (PWt / safe).astype(f32), PWt)
    Pcn = np.where(pos, (Pcsf / safe).astype(f32), Pcsf)
    Pcn = np.where(pos & (Pcn.astype(np.float64) > 0.1), f32(1), Pcn)
    keep = pos & (Pcn < 1)
    den = (Pcn + PWn + PGn).astype(f32)
    pw = np.where(keep, PWn / np.where(keep, den, f32(1)), f32(0)).astype(f32)
    pg = np.where(keep, PGn / np.where(keep, den, f32(1)), f32(0)).astype(f32)
    # tumour (Real arithmetic)
    t = np.asarray(tumor_ic, f32)
    px = (x - np.float64(t[0])).astype(f32)
    py = (y - np.float64(t[1])).astype(f32)
    pz = (z - np.float64(t[2])).astype(f32)
    dist = np.sqrt(px * px + py * py + pz * pz, dtype=f32)
    iw = f32(1.0) / (f32(2.0) * (f32(1.0) / f32(128)))
    psi = ((dist - f32(0.005)) * iw).astype(f32)
    smooth = (0.5 * ((1 - psi) - np.sin(np.pi * psi.astype(np.float64)) / np.pi)).astype(f32)
    phi = np.where((psi < -1) & ((PGn.astype(np.float64) > 0.001) | (PWn.astype(np.float64) > 0.001)), f32(1.0),
                   np.where((psi >= -1) & (psi <= 1) & ((PGn > 0) | (PWn > 0)), smooth, f32(0)))
    return phi.astype(f32), pw, pg
