"""CPU checks of the product's host code through the C-ABI (no GPU needed):
the library loads and exports every entry point of include/gq_api.h; the
host topology (gq_topo_*) plans refine/compress exactly as the oracle
topology does; its ghost expansion equals the oracle's term for term."""
import os
import re

import numpy as np
import pytest

import gliomauq_amd as gq
from gliomauq_amd import _lib
from oracle import oracle as O
from oracle import topology as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "gq_api.h")).read()
    declared = set(re.findall(r"\b(gq_[a-z0-9_]+)\s*\(", hdr))
    L = _lib.load()
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert declared == set(_lib.EXPORTS)


def test_status_strings():
    L = _lib.load()
    for s in range(8):
        assert L.gq_status_string(s)
    assert L.gq_status_string(0) == b"ok"


def keyset(lv, ix):
    return {(int(l),) + tuple(int(c) for c in i) for l, i in zip(lv, ix)}


def oracle_ids_for(topo, keys):
    return {topo.leaves[k] for k in keys}


def product_ids_for(lv, ix, ids, keys):
    m = {(int(l),) + tuple(int(c) for c in i): int(b) for l, i, b in zip(lv, ix, ids)}
    return {m[k] for k in keys}


@pytest.mark.parametrize("order", [2, 4])
def test_topology_matches_oracle(order):
    rng = np.random.default_rng(order)
    prod = gq.Topology(4, 3, order)
    orc = T.Topology(4, 3, B=16, order=order)
    for step in range(8):
        lv, ix, ids = prod.leaves()
        keys = sorted(keyset(lv, ix))
        assert set(keys) == set(orc.leaves)
        pick = [keys[i] for i in rng.choice(len(keys), size=max(1, len(keys) // 6), replace=False)]
        if step % 2 == 0:
            n1 = prod.refine(product_ids_for(lv, ix, ids, pick))
            parents = orc.plan_refine(oracle_ids_for(orc, pick))
            # apply to oracle topology (no data)
            for p in sorted(parents):
                orc.leaves.pop(p)
                for c in range(8):
                    orc.leaves[(p[0] + 1, 2 * p[1] + (c & 1), 2 * p[2] + ((c >> 1) & 1), 2 * p[3] + (c >> 2))] = \
                        orc._new_id()
            assert n1 == len(parents)
        else:
            # flag everything at the deepest levels -> compress where allowed
            lvmax = max(k[0] for k in keys)
            pick = [k for k in keys if k[0] >= lvmax - 1 and rng.random() < 0.9]
            n1 = prod.compress(product_ids_for(lv, ix, ids, pick))
            plan = orc.plan_compress(oracle_ids_for(orc, pick))
            for p in plan:
                for c in range(8):
                    orc.leaves.pop((p[0] + 1, 2 * p[1] + (c & 1), 2 * p[2] + ((c >> 1) & 1), 2 * p[3] + (c >> 2)))
                orc.leaves[p] = orc._new_id()
            assert n1 == len(plan)
    lv, ix, ids = prod.leaves()
    assert keyset(lv, ix) == set(orc.leaves)


@pytest.mark.parametrize("order", [2, 4])
def test_product_ghost_expansion_equals_oracle(order):
    prod = gq.Topology(4, 4, order)
    # refine twice around one point: levels 2, 3 and 4
    for p in [(0.3, 0.6, 0.45)]:
        for _ in range(2):
            lv, ix, ids = prod.leaves()
            sel = {int(b) for l, i, b in zip(lv, ix, ids)
                   if all(i[a] <= p[a] * (1 << l) < i[a] + 1 for a in range(3))}
            prod.refine(sel)
    lv, ix, ids = prod.leaves()
    assert set(lv.tolist()) == {2, 3, 4}
    g = O.LeafGrid(lv, ix, 16, 4, order)
    s0, s1 = (-2, 4) if order == 4 else (-1, 2)
    rng = np.random.default_rng(1)
    targets = rng.choice(len(lv), size=6, replace=False)
    n = 0
    for t in targets:
        pts = [(x, y, z) for z in range(s0, 16 + s1 - 1, 3) for y in range(s0, 16 + s1 - 1, 2)
               for x in range(s0, 16 + s1 - 1) if not (0 <= x < 16 and 0 <= y < 16 and 0 <= z < 16)]
        for p in pts:
            a = prod.point_terms(int(t), p)
            b = g.point_terms(int(t), p)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), (t, p)
            n += 1
    assert n > 1000


def test_solver_params_round_as_reference():
    """Glioma_HG_UQ::run reads Dw, rho into Real = float and stores Dw/L^2 and
    0.1*Dw as float (Glioma_HG_UQ.cpp:519-533); dt is formed in double from
    those floats (:538-539).  The Python mirror hands the same values to the
    step as gq_brain does."""
    from types import SimpleNamespace
    s = gq.Glioma_HG_UQ(SimpleNamespace(bpd=16), 17.0, cfl=0.9)
    for Dw_cm, rho in ((0.0013, 0.025), (0.002385557629465, 0.0175151508213), (1e-4 / 3, 0.1)):
        Dw, Dg, r, dt = s.params(Dw_cm, rho)
        for v in (Dw, Dg, r):
            assert float(np.float32(v)) == v
        assert Dw == float(np.float32(np.float64(np.float32(Dw_cm)) / 289.0))
        assert Dg == float(np.float32(0.1 * Dw))
        assert r == float(np.float32(rho))
        h = 1.0 / 256
        assert dt == 0.9 * h * h / (6.0 * max(Dw, Dg))
    # the rounding matters: the double quotient differs from the float one
    assert s.params(0.0013, 0.025)[0] != 0.0013 / 289.0
