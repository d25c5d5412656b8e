"""What the trained surrogate gains as the first stage of delayed acceptance.

Delayed-acceptance RWM (gliomauq_amd.sampler.da_rwm, Christen & Fox 2005)
screens a proposal t' from the state t with the surrogate posterior s
(alpha1 = min(1, exp(s(t') - s(t)))) and solves the full model f only for
the survivors (alpha2 = min(1, exp((f(t') - f(t)) - (s(t') - s(t))))).  Over a
set of proposals the expected full solves per accepted move are
E[alpha1] / E[alpha1 alpha2]; plain RWM, the method without the network,
needs 1 / E[min(1, exp(f(t') - f(t)))].  The test trains the surrogate on
fields the batched solver generates (bf16 MFMA forward/backward kernels,
Adam), then evaluates every proposal of a seeded set with the full solver
and with the trained and an untrained network.

State: a burn-in state one training half-width away from the truth in
every parameter, proposals at 0.5 and 0.1 of that half-width (64^3 grid,
500 steps, flip-rate likelihood).  Measured on MI355X: trained Spearman
rank correlation with the full log-posterior differences 0.94/0.97, full
solves per accepted move 1.47x/1.58x fewer than plain RWM; the untrained
network anti-correlates (-0.59/-0.72) and needs 2-3x more.  Near the mode the
posterior is too narrow for the network's field error and the first stage
gains nothing (the benchmark's surrogate workload shows that side).

Bars: trained Spearman >= 0.85, gain >= 1.2 and above the untrained gain.
"""
import numpy as np
import pytest
from scipy.stats import spearmanr

pytestmark = pytest.mark.gpu

L_CM = 17.0


def _lme(x):
    m = np.max(x)
    return m + np.log(np.mean(np.exp(x - m)))


def test_trained_surrogate_cuts_full_solves_per_accepted_move():
    import gliomauq_amd as gq
    from gliomauq_amd import anatomy as A
    from gliomauq_amd import sampler as SM
    from gliomauq_amd import surrogate as SU

    n, steps, B = 64, 500, 64
    rng = np.random.default_rng(5)
    pwm, pgm, _, c = A.config_inputs(n, L_CM, 0)
    h = 1.0 / n
    dt = 0.9 * h * h / (6 * 1.5e-2 / L_CM ** 2)
    tt = np.array([np.log(0.0013 * 1.25), np.log(0.025 * 0.8) + 1.0, c[2] + 0.3, c[1] - 0.4, c[0] + 0.2, 0.1])
    lo = np.array([np.log(2e-4), np.log(2e-3), c[2] - 6, c[1] - 6, c[0] - 6, 0.02])
    hi = np.array([np.log(1.5e-2), np.log(0.2), c[2] + 6, c[1] + 6, c[0] + 6, 0.3])
    span = np.array([0.3, 0.3, 2.0, 2.0, 2.0, 0.05])

    ctx = gq.Context(0)
    truth = gq.Ensemble(ctx, 1, n // 16, 2, 4)
    truth.write_volume("pw", pwm)
    truth.write_volume("pg", pgm)
    truth.ic_gaussian(np.array([[*tt[2:5], tt[5], 2.0]]))
    Dt = np.exp(tt[0]) / L_CM ** 2
    truth.run(np.array([[Dt, 0.1 * Dt, np.exp(tt[1]), dt]]), [steps], adapt_every=0)
    obs = A.observations(truth.sample(0).dump_uniform(), seed=0)
    del truth
    ctx.close()

    post = SM.Posterior(gq.Context(0), pwm, pgm, obs, steps, dt, lo, hi, batch=B)

    # training fields from the batched solver over the box tt +- span
    ntr = 256
    ths = np.clip(tt + span * rng.uniform(-1, 1, (ntr, 6)), lo, hi)
    fields = []
    for i in range(0, ntr, B):
        m, _, prm, ic = post._inputs(ths[i:i + B])
        post.ens.ic_gaussian(ic)
        st = np.zeros(B, np.int32)
        st[:m] = steps
        post.ens.run(prm, st, adapt_every=0)
        fields.extend(post.ens.sample(s).dump_uniform().astype(np.float32) for s in range(m))
    fields = np.array(fields)

    mu, sd = tt[[0, 1, 5]], span[[0, 1, 5]]
    sur = SU.Surrogate(batch=8, seed=0)
    sur.set_anatomy(pwm, pgm)
    spost = SM.SurrogatePosterior(sur, obs, lo, hi, ratio=1, norm_mu=mu, norm_sd=sd)
    untr = SU.Surrogate(batch=8, seed=1)
    untr.set_anatomy(pwm, pgm)
    spost_u = SM.SurrogatePosterior(untr, obs, lo, hi, ratio=1, norm_mu=mu, norm_sd=sd)
    P, U0 = spost.inputs(ths)
    epochs = 8
    first = last = None
    for ep in range(epochs):
        perm = rng.permutation(ntr)
        lr = 3e-4 * 0.5 ** (ep // 2)
        ls = [sur.train_step(P[perm[i:i + 8]], U0[perm[i:i + 8]], fields[perm[i:i + 8]], lr)
              for i in range(0, ntr, 8)]
        first = np.mean(ls) if first is None else first
        last = np.mean(ls)
    print("training loss %.4e -> %.4e" % (first, last))
    assert last < 0.5 * first

    def s_lp(sp, th):
        return np.concatenate([sp.logpost(th[k:k + 8]) for k in range(0, len(th), 8)])

    def f_lp(th):
        return np.concatenate([post.logpost(th[k:k + B]) for k in range(0, len(th), B)])

    thc = np.clip(tt + span * np.array([1, -1, 1, 1, -1, 1]), lo, hi)
    f0, s0, u0 = f_lp(thc[None])[0], s_lp(spost, thc[None])[0], s_lp(spost_u, thc[None])[0]
    for scale in (0.5, 0.1):
        prop = np.clip(thc + scale * span * rng.standard_normal((256, 6)), lo, hi)
        df = f_lp(prop) - f0
        assert np.all(np.isfinite(df))
        plain = np.exp(-_lme(np.minimum(0.0, df)))
        res = {}
        for name, d in (("trained", s_lp(spost, prop) - s0), ("untrained", s_lp(spost_u, prop) - u0)):
            a1 = np.minimum(0.0, d)
            a12 = a1 + np.minimum(0.0, df - d)
            da = np.exp(_lme(a1) - _lme(a12))
            res[name] = (float(spearmanr(d, df).correlation), plain / da)
        print("scale %.2f: plain %.3f full solves per accepted move; trained rho %.3f gain %.2fx; "
              "untrained rho %.3f gain %.2fx" % (scale, plain, *res["trained"], *res["untrained"]))
        assert res["trained"][0] >= 0.85
        assert res["trained"][1] >= 1.2
        assert res["trained"][1] > res["untrained"][1]
    sur.close()
    untr.close()
