"""Parity at the full size of the wavelet-adaptive single solve: the
256^3-equivalent grid (16^3 blocks at levels <= 4, Interp4, refine/compress
tolerances 1e-3/1e-4 every 10 steps) on the synthetic anatomy, after 30
steps with three adaptations.  On the GPU's FP32 snapshot of that grid:

  * FWT details of every leaf bit-identical to the FP64 oracle's;
  * the uniform 256^3 dump bit-identical to the oracle's replication;
  * 10 further FK steps within relative L2 and max-abs 1e-5 of the FP64
    oracle stepping the same grid (ghost values across levels included);
  * the flip-rate log-likelihood of the 256^3 field within 1e-10 (relative)
    of the oracle's long-double sums;
  * every leaf's 26 neighbours within one level (the refinement invariant),
    and phi finite in [0, 1].
"""
import numpy as np
import pytest

import gliomauq_amd as gq
from gliomauq_amd import anatomy as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L_CM = 17.0


def test_full_size_adaptive_solve_matches_oracle():
    n, max_level, order = 256, 4, 4
    pwm, pgm, u0, _ = A.config_inputs(n, L_CM, 0)
    ctx = gq.Context(0)
    g = gq.Grid(ctx, n // 16, max_level, order)
    g.sample_volume("phi", u0)
    g.sample_volume("pw", pwm)
    g.sample_volume("pg", pgm)
    s = gq.Glioma_HG_UQ(g, L_CM, cfl=0.9, adaptive=True, tol_refine=1e-3, tol_compress=1e-4, adapt_every=10)
    Dw, Dg, rho, dt = s.params(0.0013, 0.025)
    for _ in range(3):
        gq.fk_steps(g, Dw, Dg, rho, dt, 10)
        s._adapt()
    lv, ix, _ = g.levels_indices()
    print("full-size grid: %d leaves, levels %s" % (len(lv), sorted(set(lv.tolist()))))
    assert len(set(lv.tolist())) >= 2

    # refinement invariant: the 26 neighbours of every leaf are within one level
    lut = np.full((16,) * 3, -1, np.int64)
    for k, (l, i) in enumerate(zip(lv, ix)):
        s_ = 1 << (max_level - l)
        lut[i[2] * s_:(i[2] + 1) * s_, i[1] * s_:(i[1] + 1) * s_, i[0] * s_:(i[0] + 1) * s_] = l
    assert (lut >= 0).all()
    for d in [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if (a, b, c) != (0, 0, 0)]:
        assert np.abs(lut - np.roll(lut, d, axis=(0, 1, 2))).max() <= 1

    phi, pw, pg = g.read("phi"), g.read("pw"), g.read("pg")
    assert np.isfinite(phi).all() and phi.min() >= 0 and phi.max() <= 1
    og = O.LeafGrid(lv, ix, 16, max_level, order)

    det = gq.BlockFWT.details(g)
    ref = og.fwt_details(phi.astype(np.float64))
    assert np.array_equal(det, ref), np.max(np.abs(det - ref))

    u = g.dump_uniform()
    assert np.array_equal(u, og.dump_uniform(phi.astype(np.float64), n))

    t1, t2, pet = A.observations(u, seed=3)
    lik = gq.FlipLikelihood(g, t1, t2, pet).loglik()
    l_ref, parts = O.lik_flip(u, t1, t2, pet)
    assert abs(lik[0] - l_ref) <= 1e-10 * abs(l_ref)

    steps = 10
    gq.fk_steps(g, Dw, Dg, rho, dt, steps)
    got = g.read("phi").astype(np.float64)
    exp = og.fk_steps(phi.astype(np.float64), pw.astype(np.float64), pg.astype(np.float64), Dw, Dg, rho, dt, steps,
                      threads=8)
    e2 = float(np.sqrt(np.mean((got - exp) ** 2)) / np.sqrt(np.mean(exp ** 2)))
    einf = float(np.max(np.abs(got - exp)))
    print("full-size FK after %d steps: rel L2 %.3e, max abs %.3e; loglik %.6f" % (steps, e2, einf, lik[0]))
    assert e2 <= 1e-5 and einf <= 1e-5
    ctx.close()
