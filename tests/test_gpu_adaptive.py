"""GPU parity of the adaptive path (ghosts across levels, FWT details,
refine/compress decisions, split/collapse, dump) against the CPU oracle.

Bars: FWT details and refine/compress decisions bit-identical on identical
FP32 snapshots; split/collapse data equal to the oracle's FP64 transfer
rounded to FP32 (max-abs <= 2e-7 relative to the field scale); the FK step
on a multi-level grid within relative L2 1e-5 of the FP64 oracle.
"""
import numpy as np
import pytest

import gliomauq_amd as gq
from oracle import oracle as O
from oracle import topology as T

pytestmark = pytest.mark.gpu


def smooth_fields(x, y, z):
    phi = 0.8 * np.exp(-((x - 0.41) ** 2 + (y - 0.38) ** 2 + (z - 0.45) ** 2) / 0.006)
    pw = 0.5 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * (y + z))
    pg = 1.0 - pw
    csf = ((x - 0.6) ** 2 + (y - 0.55) ** 2 + (z - 0.5) ** 2) < 0.003
    pw = np.where(csf, 0.0, pw)
    pg = np.where(csf, 0.0, pg)
    return phi, pw, pg


def make_grid(ctx, bpd, max_level, order, refine_points=()):
    g = gq.Grid(ctx, bpd, max_level, order)
    for name, k in (("phi", 0), ("pw", 1), ("pg", 2)):
        g.sample(name, lambda x, y, z, k=k: smooth_fields(x, y, z)[k])
    for p in refine_points:
        lv, ix, ids = g.levels_indices()
        sel = {int(b) for l, i, b in zip(lv, ix, ids) if all(i[a] <= p[a] * (1 << l) < i[a] + 1 for a in range(3))}
        g.refine(sel)
    return g


def snapshot(g):
    lv, ix, ids = g.levels_indices()
    return lv, ix, ids, g.read("phi"), g.read("pw"), g.read("pg")


def keyed(lv, ix, arr):
    return {(int(l),) + tuple(int(c) for c in i): a for l, i, a in zip(lv, ix, arr)}


@pytest.mark.parametrize("order", [2, 4])
def test_fwt_details_bit_identical(ctx, order):
    g = make_grid(ctx, 4, 4, order, [(0.41, 0.38, 0.45), (0.41, 0.38, 0.45)])
    lv, ix, ids, phi, pw, pg = snapshot(g)
    assert len(set(lv.tolist())) >= 2
    det = gq.BlockFWT.details(g)
    og = O.LeafGrid(lv, ix, 16, 4, order)
    ref = og.fwt_details(phi.astype(np.float64))
    assert np.array_equal(det, ref), np.max(np.abs(det - ref))
    assert (det > 0).any()


@pytest.mark.parametrize("order", [2, 4])
def test_split_and_collapse_match_oracle(ctx, order):
    g = make_grid(ctx, 4, 3, order)
    lv, ix, ids, phi, pw, pg = snapshot(g)
    topo = T.Topology(4, 3, B=16, order=order)
    af = T.AdaptiveField(topo)
    kid = {k: i for k, i in topo.leaves.items()}
    for l, i, a, b, c in zip(lv, ix, phi, pw, pg):
        k = (int(l),) + tuple(int(v) for v in i)
        af.data[kid[k]] = {"phi": a.astype(np.float64), "pw": b.astype(np.float64), "pg": c.astype(np.float64)}
    # refine the block holding the tumour centre
    pick = [(2, 1, 1, 1)]
    g_ids = {int(bid) for l, i, bid in zip(lv, ix, ids) if (int(l),) + tuple(int(v) for v in i) in pick}
    n_gpu = g.refine(g_ids)
    n_orc = af.refine({kid[k] for k in pick})
    assert n_gpu == n_orc
    lv2, ix2, ids2, phi2, pw2, pg2 = snapshot(g)
    assert {(int(l),) + tuple(int(v) for v in i) for l, i in zip(lv2, ix2)} == set(topo.leaves)
    for name, arr in (("phi", phi2), ("pw", pw2), ("pg", pg2)):
        mine = keyed(lv2, ix2, arr)
        for k, bid in topo.leaves.items():
            ref = af.data[bid][name]
            assert np.max(np.abs(mine[k] - ref)) <= 2e-7 * max(1.0, np.abs(ref).max()), (name, k)
    # collapse the children again (all flagged): injection of even points
    lvc, ixc, idsc = g.levels_indices()
    fine = {(int(l),) + tuple(int(v) for v in i): int(b) for l, i, b in zip(lvc, ixc, idsc) if l == 3}
    n_c = g.compress(set(fine.values()))
    n_o = af.compress({topo.leaves[k] for k in fine})
    assert n_c == n_o == 1
    lv3, ix3, _, phi3, pw3, pg3 = snapshot(g)
    assert {(int(l),) + tuple(int(v) for v in i) for l, i in zip(lv3, ix3)} == set(topo.leaves)
    for name, arr in (("phi", phi3), ("pw", pw3), ("pg", pg3)):
        mine = keyed(lv3, ix3, arr)
        for k, bid in topo.leaves.items():
            assert np.array_equal(mine[k].astype(np.float64), af.data[bid][name].astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("order", [2, 4])
def test_adaptive_step_parity(ctx, order):
    g = make_grid(ctx, 4, 4, order, [(0.41, 0.38, 0.45), (0.41, 0.38, 0.45)])
    lv, ix, ids, phi, pw, pg = snapshot(g)
    assert set(lv.tolist()) == {2, 3, 4}
    og = O.LeafGrid(lv, ix, 16, 4, order)
    Dw = 2e-5
    Dg = 0.1 * Dw
    rho = 0.3
    h = 1.0 / (16 * 16)
    dt = 0.9 * h * h / (6 * Dw)
    steps = 25
    gq.fk_steps(g, Dw, Dg, rho, dt, steps)
    phi_gpu = g.read("phi").astype(np.float64)
    ref = og.fk_steps(phi.astype(np.float64), pw.astype(np.float64), pg.astype(np.float64), Dw, Dg, rho, dt, steps)
    e = float(np.sqrt(np.mean((phi_gpu - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    einf = float(np.max(np.abs(phi_gpu - ref)))
    print("adaptive order %d: rel L2 %.3e Linf %.3e" % (order, e, einf))
    assert e <= 1e-5 and einf <= 1e-5
    assert np.count_nonzero(((phi_gpu == 0) != (ref == 0)) & (np.maximum(phi_gpu, ref) > 1e-30)) == 0


@pytest.mark.parametrize("order", [4])
def test_auto_refine_compress_decisions_match_oracle(ctx, order):
    """Science::AutomaticCompression then AutomaticRefinement from the same
    snapshot: identical leaf sets (the decisions are bit-exact functions of
    the FP32 snapshot)."""
    g = make_grid(ctx, 8, 3, order)
    lv, ix, ids, phi, pw, pg = snapshot(g)
    topo = T.Topology(8, 3, B=16, order=order)
    af = T.AdaptiveField(topo)
    for l, i, a, b, c in zip(lv, ix, phi, pw, pg):
        k = (int(l),) + tuple(int(v) for v in i)
        af.data[topo.leaves[k]] = {"phi": a.astype(np.float64), "pw": b.astype(np.float64),
                                   "pg": c.astype(np.float64)}
    tol_c, tol_r = 1e-4, 1e-3
    n_gpu = gq.Science.AutomaticCompression(g, tol_c, -1)
    n_orc = af.automatic_compression(tol_c, -1)
    lv2, ix2, _ = g.levels_indices()
    assert {(int(l),) + tuple(int(v) for v in i) for l, i in zip(lv2, ix2)} == set(topo.leaves)
    assert n_gpu == n_orc and n_gpu > 0
    # refinement decisions on the compressed grid; the oracle continues from
    # the GPU's FP32 snapshot so both decide on identical data
    lv2, ix2, ids2, phi2, pw2, pg2 = snapshot(g)
    for l, i, a, b, c in zip(lv2, ix2, phi2, pw2, pg2):
        k = (int(l),) + tuple(int(v) for v in i)
        af.data[topo.leaves[k]] = {"phi": a.astype(np.float64), "pw": b.astype(np.float64),
                                   "pg": c.astype(np.float64)}
    n_gpu = gq.Science.AutomaticRefinement(g, tol_r, 3, 1)
    n_orc = af.automatic_refinement(tol_r, 3, 1)
    lv3, ix3, _ = g.levels_indices()
    assert {(int(l),) + tuple(int(v) for v in i) for l, i in zip(lv3, ix3)} == set(topo.leaves)
    assert n_gpu == n_orc


def test_dump_uniform_matches_oracle(ctx):
    g = make_grid(ctx, 8, 3, 4)
    lv, ix, ids = g.levels_indices()
    # compress a corner region twice: levels 3, 2 and 1 (r = 1, 2, 4)
    for _ in range(2):
        lv, ix, ids = g.levels_indices()
        sel = {int(b) for l, i, b in zip(lv, ix, ids) if all(i[a] * 2.0 ** -l >= 0.5 for a in range(3))}
        g.compress(sel)
    lv, ix, ids, phi, pw, pg = snapshot(g)
    assert {1, 2, 3} <= set(lv.tolist()) or {2, 3} <= set(lv.tolist())
    u = g.dump_uniform()
    og = O.LeafGrid(lv, ix, 16, 3, 4)
    ref = og.dump_uniform(phi.astype(np.float64), 128)
    assert np.array_equal(u, ref)
