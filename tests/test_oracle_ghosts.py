"""Pins of the oracle's ghost expansion (og_* in oracle/fk_oracle.c).

* equal, as merged weight maps, to the reference's own MRAG_BBInfoCreator
  (compiled from source, oracle/_ref) on balanced two/three-level toy grids;
* equal to committed golden lists made by that builder (tests/golden);
* polynomial exactness: Interp4 prolongation reproduces cubics and Interp2
  reproduces linear functions at every ghost of a refined block.
"""
import os

import numpy as np
import pytest

from helpers import balanced_leaves, leaf_arrays
from oracle import oracle as O
from oracle import ref as R

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def halo_points(B, s0, s1):
    for z in range(s0, B + s1 - 1):
        for y in range(s0, B + s1 - 1):
            for x in range(s0, B + s1 - 1):
                if 0 <= x < B and 0 <= y < B and 0 <= z < B:
                    continue
                yield (x, y, z)


def merged(leaf, cell, w):
    d = {}
    for a, b, c in zip(leaf, cell, w):
        d[(int(a), int(b))] = d.get((int(a), int(b)), 0.0) + float(c)
    return {k: v for k, v in d.items() if v != 0.0}


def compare_with_reference(lv, ix, order, B, targets):
    s0, s1 = (-2, 4) if order == 4 else (-1, 2)
    g = O.LeafGrid(lv, ix, B, int(lv.max()), order)
    n = 0
    for t in targets:
        off, leaf, cell, w = R.ghost_terms(lv, ix, t, order, B, s0, s1)
        for k, p in enumerate(halo_points(B, s0, s1)):
            a, b = off[k], off[k + 1]
            ref = merged(leaf[a:b], cell[a:b], w[a:b])
            ol, oc, ow = g.point_terms(t, p)
            assert merged(ol, oc, ow) == ref, (t, p)
            n += 1
    return n


need_ref = pytest.mark.skipif(not R.available("libgqref_ghosts.so"), reason="oracle/_ref not built")


@need_ref
@pytest.mark.parametrize("order", [2, 4])
def test_ghosts_match_reference_two_levels(order):
    lv, ix = leaf_arrays(balanced_leaves(1, [(0.3, 0.3, 0.3)], 2))
    assert compare_with_reference(lv, ix, order, 8, range(len(lv))) > 0


@need_ref
@pytest.mark.parametrize("order", [2, 4])
def test_ghosts_match_reference_three_levels(order):
    lv, ix = leaf_arrays(balanced_leaves(2, [(0.40, 0.45, 0.5), (0.9, 0.1, 0.2)], 4))
    assert set(lv.tolist()) == {2, 3, 4}
    rng = np.random.default_rng(0)
    targets = sorted(set(rng.choice(len(lv), 12, replace=False).tolist()) | {int(np.argmax(lv)), int(np.argmin(lv))})
    assert compare_with_reference(lv, ix, order, 8, targets) > 0


@pytest.mark.parametrize("order", [2, 4])
def test_ghosts_golden(order):
    gd = np.load(os.path.join(GOLD, "ghosts.npz"))
    lv, ix = gd["levels"], gd["indices"]
    off, leaf, cell, w = gd["off%d" % order], gd["leaf%d" % order], gd["cell%d" % order], gd["w%d" % order]
    s0, s1 = (-2, 4) if order == 4 else (-1, 2)
    g = O.LeafGrid(lv, ix, 8, int(lv.max()), order)
    k = 0
    for t in range(len(lv)):
        for p in halo_points(8, s0, s1):
            a, b = off[k], off[k + 1]
            ref = merged(leaf[a:b], cell[a:b], w[a:b])
            ol, oc, ow = g.point_terms(t, p)
            assert merged(ol, oc, ow) == ref, (t, p)
            k += 1
    assert k == len(off) - 1


@pytest.mark.parametrize("order,deg", [(4, 3), (2, 1)])
def test_prolongation_reproduces_polynomials(order, deg):
    """Fine-block ghosts inside a coarse neighbour equal the polynomial."""
    B = 8
    lv, ix = leaf_arrays(balanced_leaves(1, [(0.55, 0.55, 0.55)], 2))
    g = O.LeafGrid(lv, ix, B, 2, order)
    rng = np.random.default_rng(4)
    cz = rng.standard_normal(deg + 1)
    cy = rng.standard_normal(deg + 1)
    cx = rng.standard_normal(deg + 1)

    def f(x, y, z):
        return np.polyval(cx, x) + np.polyval(cy, y) * np.polyval(cz, z)

    data = np.zeros((len(lv), B, B, B))
    for i, (l, (a, b, c)) in enumerate(zip(lv, ix)):
        h = 2.0 ** -l / B
        r = np.arange(B) * h
        z, y, x = np.meshgrid(c * 2.0 ** -l + r, b * 2.0 ** -l + r, a * 2.0 ** -l + r, indexing="ij")
        data[i] = f(x, y, z)
    # a level-2 block whose lab reaches into level-1 neighbours, away from
    # the periodic wrap (the polynomial is not periodic)
    t = next(i for i, (l, idx) in enumerate(zip(lv, ix)) if l == 2 and tuple(idx) == (2, 2, 2))
    s0, s1 = (-2, 4) if order == 4 else (-1, 2)
    lab = g.fill_lab(data, t, s0, s1)
    h = 0.25 / B
    r = np.arange(s0, B + s1 - 1) * h
    z, y, x = np.meshgrid(0.5 + r, 0.5 + r, 0.5 + r, indexing="ij")
    np.testing.assert_allclose(lab, f(x, y, z), rtol=0, atol=1e-12)
