"""Writes tests/golden/adaptive_solve256.npz: the FP64 oracle's own end-to-end
run of the wavelet-adaptive solve (Glioma_HG_UQ::run, Glioma/Glioma_HG_UQ.cpp:
544-575) on the 256^3-equivalent synthetic-anatomy workload -- seed 0, 17 cm
cube, Dw = 0.0013 cm^2/day, rho = 0.025/day, dt = 0.9 CFL of the finest level,
T = 300 days (590 steps), Interp4, refine/compress tolerances 1e-3/1e-4 after
every 10 steps, levels <= 4, final refinement -- with the oracle making its own
refine/compress decisions on its FP64 field.

Stored: the leaf count after every adaptation, the final leaves (level,
index), and the final uniform dump (Glioma_HG_UQ::_dumpUQoutput) as its
voxels above TRUNC = 1e-6 (flat indices as increments, values rounded to
float32); the voxels left out hold the tail of the front (0 < u <= 1e-6), so
the stored field differs from the oracle's by at most 1e-6 anywhere.  The GPU test compares
its own FP32 run with these.

Run from the repository root (about 10 minutes on 8 cores):
    python tests/golden/make_adaptive_golden.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gliomauq_amd import anatomy as A  # noqa: E402
from oracle import topology as T  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adaptive_solve256.npz")
TRUNC = 1e-6
L_CM = 17.0


def params(n, Dw_cm=0.0013, rho=0.025, cfl=0.9):
    """(Dw, Dg, rho, dt) as Glioma_HG_UQ::run forms them (float storage)."""
    f32 = np.float32
    Dw = float(f32(float(f32(Dw_cm)) / (L_CM * L_CM)))
    Dg = float(f32(0.1 * Dw))
    h = 1.0 / n
    return Dw, Dg, float(f32(rho)), cfl * h * h / (6.0 * max(Dw, Dg))


def n_steps_until(T_days, dt):
    t, n = 0.0, 0
    while t <= T_days:
        n += 1
        t += dt
    return n


def main(n=256, threads=8):
    pwm, pgm, u0, _ = A.config_inputs(n, L_CM, 0)
    bpd = n // 16
    max_level = int(round(np.log2(bpd)))
    Dw, Dg, rho, dt = params(n)
    steps = n_steps_until(300.0, dt)
    topo = T.Topology(bpd, max_level, B=16, order=4)
    fld = T.AdaptiveField(topo)
    vols = {"phi": u0, "pw": pwm, "pg": pgm}
    for name, v in vols.items():
        fld.set_from_volume(name, lambda l, org, h, v=v: v[int(round(org[2] * n)):int(round(org[2] * n)) + 16,
                                                          int(round(org[1] * n)):int(round(org[1] * n)) + 16,
                                                          int(round(org[0] * n)):int(round(org[0] * n)) + 16])
    t0 = time.time()
    hist = T.run_adaptive(fld, Dw, Dg, rho, dt, steps, 10, 1e-3, 1e-4, max_level, threads,
                          log=lambda m: print("%7.1f s  %s" % (time.time() - t0, m), flush=True))
    keys, g = topo.leaf_grid()
    dump = g.dump_uniform(fld.stacked("phi", keys).reshape(len(keys), -1), n).reshape(-1)
    nz = np.flatnonzero(dump > TRUNC)
    np.savez_compressed(OUT, n=n, trunc=TRUNC, steps=steps, dt=dt, Dw=Dw, Dg=Dg, rho=rho, history=np.array(hist, np.int32),
                        levels=np.array([k[0] for k in keys], np.int8),
                        index=np.array([k[1:] for k in keys], np.int16),
                        dump_step=np.diff(nz, prepend=0).astype(np.uint32), dump_val=dump[nz].astype(np.float32))
    print("wrote %s: %d steps, %d adaptations, %d final leaves, %d voxels above 1e-6, %.0f s"
          % (OUT, steps, len(hist) - 1, len(keys), len(nz), time.time() - t0))


if __name__ == "__main__":
    main(*(int(a) for a in sys.argv[1:]))
