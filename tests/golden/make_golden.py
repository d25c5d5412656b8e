"""Writes the golden fixtures of tests/golden from the reference's own code
compiled from source into oracle/_ref (`make -C oracle ref`).

  fk_operator.npz  Glioma_ReactionDiffusionOperator<3> + UpdateTumor<3> on a
                   seeded 8^3 block lab (Real=double and Real=float)
  ghosts.npz       MRAG_BBInfoCreator instruction lists (flattened, merged per
                   source point) for every block of a seeded 2-level hierarchy,
                   Interp4 stencil [-2,4) and Interp2 stencil [-1,2), B=8
  likelihood.npz   the reference likelihood program on a seeded 16^3 input
                   (selective points and all-voxel modes)

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import ref as R  # noqa: E402
from helpers import balanced_leaves, leaf_arrays  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def fk():
    rng = np.random.default_rng(2024)
    B = 8
    L = B + 2
    pw = rng.random((L, L, L))
    pg = 1 - pw
    csf = rng.random((L, L, L)) < 0.1
    pw[csf] = 0
    pg[csf] = 0
    phi = rng.random((L, L, L)) * 0.6
    phi[rng.random((L, L, L)) < 0.3] = 0
    h, Dw, rho, dt = 1.0 / 128, 0.0013 / 17.0 ** 2, 0.025, 2.0353
    Dg = 0.1 * Dw
    out = dict(lab_phi=phi, lab_pw=pw, lab_pg=pg, h=h, Dw=Dw, Dg=Dg, rho=rho, dt=dt)
    for real, dtp in (("f64", np.float64), ("f32", np.float32)):
        a = [x.astype(dtp) for x in (phi, pw, pg)]
        d = R.fk_block(*a, h, Dw, Dg, rho, real=real)
        out["dphidt_" + real] = d
        out["phi1_" + real] = R.update(np.ascontiguousarray(a[0][1:-1, 1:-1, 1:-1]), d, dt, real=real)
    np.savez_compressed(os.path.join(OUT, "fk_operator.npz"), **out)


def ghost_case():
    leaves = balanced_leaves(1, [(0.3, 0.3, 0.3)], 2)
    return leaf_arrays(leaves)


def ghosts():
    lv, ix = ghost_case()
    out = dict(levels=lv, indices=ix)
    for order, (s0, s1) in ((4, (-2, 4)), (2, (-1, 2))):
        offs, lf, cl, ws = [], [], [], []
        base = 0
        for t in range(len(lv)):
            off, leaf, cell, w = R.ghost_terms(lv, ix, t, order, 8, s0, s1)
            # merge per (leaf, cell) inside each ghost, sorted by (leaf, cell)
            ng = len(off) - 1
            o2 = [0]
            for g in range(ng):
                d = {}
                for j in range(off[g], off[g + 1]):
                    d[(leaf[j], cell[j])] = d.get((leaf[j], cell[j]), 0.0) + w[j]
                items = sorted((k, v) for k, v in d.items() if v != 0.0)
                for (a, b), v in items:
                    lf.append(a)
                    cl.append(b)
                    ws.append(v)
                o2.append(o2[-1] + len(items))
            offs.append(np.array(o2) + base)
            base += o2[-1]
        out["off%d" % order] = np.concatenate([o[:-1] for o in offs] + [np.array([base])]).astype(np.int64)
        out["leaf%d" % order] = np.array(lf, np.int32)
        out["cell%d" % order] = np.array(cl, np.int32)
        out["w%d" % order] = np.array(ws, np.float64)
    np.savez_compressed(os.path.join(OUT, "ghosts.npz"), **out)


def likelihood():
    rng = np.random.default_rng(77)
    n = 16
    model = np.clip(rng.random((n, n, n)) * 1.2 - 0.1, 0, 1)
    pet = np.where(rng.random((n, n, n)) < 0.5, rng.random((n, n, n)), 0.0)
    t1 = (rng.random((n, n, n)) < 0.3).astype(np.float64)
    t2 = (rng.random((n, n, n)) < 0.5).astype(np.float64)
    pts = rng.integers(0, n, size=(200, 3)).astype(np.float64)  # rows: x, y, z
    params = [0.000361, 0.85, 0.7, 0.25, 0.09]  # Likelihood/makefile/LikelihoodInput.txt
    sel, _ = R.run_likelihood(model, pet, t1, t2, params, points_xyz=pts.T, selective=1)
    allv, _ = R.run_likelihood(model, pet, t1, t2, params, points_xyz=pts.T, selective=0)
    np.savez_compressed(os.path.join(OUT, "likelihood.npz"), model=model, pet=pet, t1=t1, t2=t2, points=pts,
                        params=np.array(params), minus_logL_selective=sel, minus_logL_all=allv)


if __name__ == "__main__":
    fk()
    ghosts()
    likelihood()
    print("golden fixtures written to", OUT)
