"""Host logic of the samplers (gliomauq_amd/sampler.py) on closed-form targets.

The samplers only see a posterior object with logpost / logpost_grad and a
batch size; a Gaussian target with known mean and covariance pins their
accept/reject arithmetic (a wrong proposal-density term or a dropped
gradient factor biases the variance), and an AR(1) chain with known
integrated autocorrelation pins the batch-means ESS.
"""
import numpy as np

from gliomauq_amd import sampler as SM


class Gauss:
    def __init__(self, mu, sd, batch=16, lo=-50, hi=50):
        self.mu = np.asarray(mu, float)
        self.sd = np.asarray(sd, float)
        self.batch = batch
        self.lo = np.full(6, lo, float)
        self.hi = np.full(6, hi, float)
        self.n = 0

    def logpost(self, th):
        th = np.atleast_2d(th)
        self.n += len(th)
        inside = np.all((th >= self.lo) & (th <= self.hi), 1)
        return np.where(inside, -0.5 * np.sum(((th - self.mu) / self.sd) ** 2, 1), -np.inf)

    def logpost_grad(self, th):
        th = np.atleast_2d(th)
        return self.logpost(th), -(th - self.mu) / self.sd ** 2


MU = np.array([0.3, -1.0, 2.0, 0.0, 1.5, -0.2])
SD = np.array([0.5, 1.0, 0.2, 2.0, 1.0, 0.1])


def test_mala_targets_gaussian():
    rng = np.random.default_rng(0)
    post = Gauss(MU, SD)
    th0 = MU + SD * rng.standard_normal((16, 6))
    out, acc, _ = SM.mala(post, th0, 3000, SD, rng, adapt_until=500)
    x = out[500:].reshape(-1, 6)
    a = acc[500:].mean()
    assert 0.45 < a < 0.7
    assert np.all(np.abs(x.mean(0) - MU) < 0.1 * SD)
    assert np.all(np.abs(x.std(0) / SD - 1) < 0.06)


def test_rwm_targets_gaussian():
    rng = np.random.default_rng(1)
    post = Gauss(MU, SD)
    out, acc, _ = SM.rwm(post, MU, 40000, SD, rng, adapt_until=4000)
    x = out[4000:]
    assert 0.15 < acc[4000:].mean() < 0.35
    assert np.all(np.abs(x.mean(0) - MU) < 0.1 * SD)
    assert np.all(np.abs(x.std(0) / SD - 1) < 0.1)


def test_ess_batch_means_ar1():
    rng = np.random.default_rng(2)
    phi = 0.9
    n = 200000
    e = rng.standard_normal(n)
    x = np.empty(n)
    x[0] = 0
    for i in range(1, n):
        x[i] = phi * x[i - 1] + e[i]
    ess = SM.ess_batch_means(x)
    expect = n * (1 - phi) / (1 + phi)  # integrated autocorrelation time (1+phi)/(1-phi)
    assert abs(ess / expect - 1) < 0.15
    assert SM.ess_batch_means(rng.standard_normal(10000)) > 7000


def test_tmcmc_reaches_posterior():
    rng = np.random.default_rng(3)
    post = Gauss(MU, SD, batch=64, lo=-6, hi=6)

    def prior(n, r):
        return r.uniform(-6, 6, size=(n, 6))

    th, ll, stages = SM.tmcmc(post, prior, 2000, rng, cov_scale=0.5, moves=5)
    assert stages[-1]["q"] == 1.0
    assert np.all(np.abs(th.mean(0) - MU) < 0.15 * SD)
    assert np.all(np.abs(th.std(0) / SD - 1) < 0.1)
