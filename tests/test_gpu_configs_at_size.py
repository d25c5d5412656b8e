"""The method extensions at their configured sizes.

* Batched ensemble: the 256 samples of the seeded prior (log-uniform Dw, rho,
  IC in white matter, T in [30, 300] days) on their own adaptive grids
  (levels 1-3, finest 128^3, Interp4 tol 1e-3/1e-4 every 10 steps), advanced
  together.  A seeded subset of samples is bit-identical to one-sample runs;
  every sample keeps the 26-neighbour level-jump invariant, phi within the
  range the final Interp4 refinement can reach from [0, 1], and a finite
  log-likelihood.
* Discrete adjoint: uniform 128^3, 2000 steps, checkpoints every 50 steps,
  two chains: gradients against the FP64 adjoint oracle (checkpointed, the
  same map in FP64; within 1e-3 of the largest component -- the FP32 and FP64
  trajectories part over 2000 steps) and against central finite differences
  of the GPU's own forward in log Dw and log rho.
* TMCMC (the reference's sampler family, makefile/Example/runHGG.sh:1-4) over
  the batched GPU likelihood: the stages reach q = 1, the log-likelihoods the
  sampler returns equal a fresh evaluation of the final population, and the
  posterior mean lies near the synthetic truth.
"""
import os

import numpy as np
import pytest

import gliomauq_amd as gq
from gliomauq_amd import anatomy as A
from gliomauq_amd import sampler as SM
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L_CM = 17.0


def level_jump_ok(levels, index, max_level):
    """Every leaf's 26 neighbours (periodic) within one level."""
    nF = 1 << max_level
    lut = np.full((nF, nF, nF), -1, np.int32)
    for l, ix in zip(levels, index):
        s = 1 << (max_level - l)
        x, y, z = (int(c) * s for c in ix)
        lut[z:z + s, y:y + s, x:x + s] = l
    assert (lut >= 0).all()
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                nb = np.roll(lut, (dz, dy, dx), axis=(0, 1, 2))
                if np.abs(nb - lut).max() > 1:
                    return False
    return True


def test_ensemble_256_samples_from_prior():
    n, S = 128, 256
    pwm, pgm, u0_gt, _ = A.config_inputs(n, L_CM, 0)
    prm, steps, ic = A.ensemble_prior(S, n, L_CM, 0, pwm=pwm)
    obs = A.observations(u0_gt, seed=0)
    ctx = gq.Context(0)
    ens = gq.Ensemble(ctx, S, n // 16, 3, 4)
    ens.write_volume("pw", pwm)
    ens.write_volume("pg", pgm)
    ens.ic_gaussian(ic)
    ens.set_observations(*obs)
    st = ens.run(prm, steps, adapt_every=10, tol_refine=1e-3, tol_compress=1e-4)
    ll = ens.loglik_flip()
    counts = ens.block_counts()
    print("256 samples: steps %d..%d, %d sample-steps, %.3g cell updates, final leaves %d..%d"
          % (steps.min(), steps.max(), steps.sum(), st["cell_updates"], counts.min(), counts.max()))
    assert np.all(np.isfinite(ll[:, 0]))
    ext = [0.0, 0.0]
    for s in range(S):
        g = ens.sample(s)
        lv, ix, _ = g.levels_indices()
        assert lv.min() >= 0 and lv.max() <= 3
        assert level_jump_ok(lv, ix, 3), s
        # every step clamps phi to [0, 1] (UpdateTumor); the final refinement (after the last step,
        # Glioma_HG_UQ.cpp:574-575) prolongs with the Interp4 filter, whose tensor-product weights
        # (L1 norm (5/4)^3) keep values of [0, 1] within [(1 - (5/4)^3)/2, (1 + (5/4)^3)/2]
        phi = g.read("phi")
        lo, hi = (1 - 1.25 ** 3) / 2, (1 + 1.25 ** 3) / 2
        assert phi.min() >= lo and phi.max() <= hi, (s, phi.min(), phi.max())
        ext[0], ext[1] = min(ext[0], float(phi.min())), max(ext[1], float(phi.max()))
    print("phi over all samples in [%.4f, %.4f]" % tuple(ext))
    # a seeded subset, each run alone on a one-sample grid with the same schedule: the same bits
    rng = np.random.default_rng(11)
    pick = sorted(set(rng.choice(S, 6, replace=False).tolist()) | {int(np.argmax(steps)), int(np.argmin(steps))})
    ref = {s: ens.sample(s).dump_uniform() for s in pick}
    lv_ref = {s: ens.sample(s).levels_indices()[:2] for s in pick}
    del ens
    for s in pick:
        one = gq.Ensemble(ctx, 1, n // 16, 3, 4)
        one.write_volume("pw", pwm)
        one.write_volume("pg", pgm)
        one.ic_gaussian(ic[s:s + 1])
        one.run(prm[s:s + 1], steps[s:s + 1], adapt_every=10, tol_refine=1e-3, tol_compress=1e-4)
        lv, ix, _ = one.sample(0).levels_indices()
        assert np.array_equal(lv, lv_ref[s][0]) and np.array_equal(ix, lv_ref[s][1]), s
        assert one.sample(0).dump_uniform().tobytes() == ref[s].tobytes(), "sample %d differs" % s
        del one
    ctx.close()


def test_adjoint_128_2000_steps_checkpointed():
    n, steps, ck = 128, 2000, 50
    pwm, pgm, u0, c = A.config_inputs(n, L_CM, 0)
    h = 1.0 / n
    dt = 0.9 * h * h / (6 * 1.5e-2 / L_CM ** 2)  # stable for every Dw of the prior
    truth = O.ic_gaussian(n, c[2] + 0.5, c[1] - 0.3, c[0] + 0.2, 0.6, 3.0)
    obs = A.observations(truth, seed=3)
    Dw = np.array([0.0013, 0.004]) / L_CM ** 2
    prm = np.stack([Dw, 0.1 * Dw, np.array([0.025, 0.05]), np.full(2, dt)], 1)
    ic = np.array([[c[2] + 0.4, c[1] + 0.1, c[0] - 0.3, 0.5, 2.0], [c[2] - 0.6, c[1] + 0.4, c[0], 0.4, 2.5]])
    ctx = gq.Context(0)

    def ensemble():
        e = gq.Ensemble(ctx, 2, n // 16, 3, 4)
        e.write_volume("pw", pwm)
        e.write_volume("pg", pgm)
        e.set_observations(*obs)
        return e

    ens = ensemble()
    ll, gr = ens.adjoint_flip(prm, steps, ckpt_every=ck, ic=ic)
    # FP64 adjoint oracle of the same map (checkpointed; 16 host threads)
    pet = obs[2]
    thr = min(16, os.cpu_count() or 1)
    for s in range(2):
        u0s = O.ic_gaussian(n, *ic[s]).astype(np.float32).astype(np.float64)
        f32 = np.float32
        uN, d0, gD, gR = O.adjoint_ckpt(u0s, pwm.astype(np.float64), pgm.astype(np.float64), h, float(f32(prm[s, 0])),
                                        float(f32(prm[s, 1])), float(f32(prm[s, 2])), dt, steps, ckpt_every=ck,
                                        threads=thr, pet=pet, sigma=0.05)
        ref = np.concatenate([[gD, gR], O.ic_grad(d0, *ic[s])])
        err = np.abs(gr[s] - ref) / np.max(np.abs(ref))
        print("chain %d: grad gpu %s\n         oracle %s\n         rel err %s" % (s, gr[s], ref, err))
        # the device differentiates its FP32 forward map, the oracle the FP64 one: over 2000 steps
        # the two trajectories part by ~1e-4 relative, and so do the gradients
        assert np.all(err <= 1e-3)
    # central differences of the GPU forward in log Dw and log rho (relative step 1e-4)
    eps = 1e-4
    for k, name in ((0, "log Dw"), (1, "log rho")):
        lp = []
        for sgn in (1, -1):
            p = prm.copy()
            if k == 0:
                p[:, 0] *= np.exp(sgn * eps)
                p[:, 1] *= np.exp(sgn * eps)
            else:
                p[:, 2] *= np.exp(sgn * eps)
            e = ensemble()
            e.ic_gaussian(ic)
            e.run(p, [steps, steps], adapt_every=0)
            lp.append(e.loglik_flip()[:, 3])
            del e
        fd = (lp[0] - lp[1]) / (2 * eps)
        rel = np.abs(fd - gr[:, k]) / np.abs(gr[:, k])
        print("%s: adjoint %s, central differences %s, rel %s" % (name, gr[:, k], fd, rel))
        assert np.all(rel <= 2e-2)
    ctx.close()


def test_tmcmc_over_batched_likelihood():
    n, steps = 32, 60
    pwm, pgm, u0, c = A.config_inputs(n, L_CM, 2)
    h = 1.0 / n
    dt = 0.9 * h * h / (6 * 1.5e-2 / L_CM ** 2)
    tt = np.array([np.log(0.01), np.log(0.1), c[2] + 0.3, c[1] - 0.4, c[0] + 0.2, 0.5])
    lo = np.array([np.log(2e-3), np.log(0.02), c[2] - 3, c[1] - 3, c[0] - 3, 0.2])
    hi = np.array([np.log(5e-2), np.log(0.4), c[2] + 3, c[1] + 3, c[0] + 3, 0.9])
    ctx = gq.Context(0)
    truth = gq.Ensemble(ctx, 1, n // 16, 1, 4)
    truth.write_volume("pw", pwm)
    truth.write_volume("pg", pgm)
    truth.ic_gaussian(np.array([[tt[2], tt[3], tt[4], tt[5], 2.0]]))
    Dt = np.exp(tt[0]) / L_CM ** 2
    truth.run(np.array([[Dt, 0.1 * Dt, np.exp(tt[1]), dt]]), [steps], adapt_every=0)
    obs = A.observations(truth.sample(0).dump_uniform(), seed=4)
    del truth
    post = SM.Posterior(gq.Context(0), pwm, pgm, obs, steps, dt, lo, hi, batch=32)
    rng = np.random.default_rng(5)

    def prior(m, r):
        return lo + (hi - lo) * r.random((m, 6))

    th, ll, stages = SM.tmcmc(post, prior, 96, rng, cov_scale=0.5, moves=2, max_stages=40)
    print("TMCMC: %d stages, q %s, acceptance %s" % (len(stages), [round(s["q"], 4) for s in stages],
                                                    [round(s["accept"], 2) for s in stages]))
    assert stages[-1]["q"] == 1.0
    fresh = np.concatenate([post.logpost(th[i:i + 32]) for i in range(0, len(th), 32)])
    assert np.array_equal(fresh, ll)
    mean, sd = th.mean(0), th.std(0)
    z = np.abs(mean - tt) / np.maximum(sd, 1e-3 * (hi - lo))
    print("posterior mean %s\n truth %s\n sd %s\n |mean-truth|/sd %s" % (mean, tt, sd, z))
    # log rho and the IC position are well determined by this data; the truth lies within the population
    assert np.all(z[[1, 2, 3, 4]] <= 4.0)
    ctx.close()
