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


def _ar1_chains(rng, n, c, phi, x0=None):
    x = np.zeros((n, c))
    x[0] = rng.standard_normal(c) if x0 is None else x0
    e = rng.standard_normal((n, c)) * np.sqrt(1 - phi * phi)
    for t in range(1, n):
        x[t] = phi * x[t - 1] + e[t]
    return x


def test_ess_chains_ar1():
    """Parallel stationary AR(1) chains: ESS = n C (1 - phi) / (1 + phi)
    (integrated autocorrelation time (1 + phi) / (1 - phi))."""
    rng = np.random.default_rng(4)
    n, c, phi = 1000, 200, 0.8
    x = _ar1_chains(rng, n, c, phi)
    exact = n * c * (1 - phi) / (1 + phi)
    assert abs(SM.ess_chains(x) / exact - 1) < 0.25
    assert abs(SM.ess_parallel(x) / exact - 1) < 0.25
    # short chains (60 samples, correlation time 39): the information in a
    # chain mean is n^2 / sum_ij phi^|i-j| samples; batch means over-count
    # (batches shorter than the correlation), the chain-means estimate does
    # not, and the parallel estimate takes the smaller
    n, c, phi = 60, 400, 0.95
    xs = _ar1_chains(rng, n, c, phi)
    k = np.arange(1, n)
    ex = c * n * n / (n + 2 * np.sum((n - k) * phi ** k))
    bm = sum(SM.ess_batch_means(xs[:, j]) for j in range(c))
    assert bm > 2 * ex
    assert abs(SM.ess_parallel(xs) / ex - 1) < 0.25
    # chains that still remember dispersed starts: lower estimate
    xd = _ar1_chains(rng, 200, 100, 0.9, x0=np.linspace(-20, 20, 100))
    assert SM.ess_chains(xd) < 0.9 * 200 * 100 * 0.1 / 1.9


class CorrGauss:
    """Correlated Gaussian: log p = -0.5 (x - mu)^T P (x - mu)."""

    def __init__(self, mu, cov, batch=64):
        self.mu = np.asarray(mu, float)
        self.cov = np.asarray(cov, float)
        self.P = np.linalg.inv(self.cov)
        self.batch = batch
        self.lo = np.full(6, -50.0)
        self.hi = np.full(6, 50.0)

    def logpost(self, th):
        d = np.atleast_2d(th) - self.mu
        return -0.5 * np.einsum("ij,jk,ik->i", d, self.P, d)

    def logpost_grad(self, th):
        d = np.atleast_2d(th) - self.mu
        return self.logpost(th), -d @ self.P


def _corr_cov():
    rng = np.random.default_rng(9)
    a = rng.standard_normal((6, 6))
    q, _ = np.linalg.qr(a)
    return q @ np.diag([1e-4, 1e-2, 0.3, 1.0, 4.0, 9.0]) @ q.T  # condition number 9e4


def test_full_covariance_preconditioner():
    """With L = chol(cov of a pilot population), MALA and batched RWM recover
    the mean and covariance of a badly conditioned correlated Gaussian; the
    proposal density of MALA uses M^-1 = (L L^T)^-1 (a diagonal stand-in
    would bias the covariance)."""
    cov = _corr_cov()
    mu = MU
    post = CorrGauss(mu, cov)
    rng = np.random.default_rng(3)
    pilot = mu + rng.multivariate_normal(np.zeros(6), cov, 400)
    L = SM.population_cholesky(pilot)
    th0 = mu + rng.multivariate_normal(np.zeros(6), cov, 64)
    for name in ("mala", "rwm"):
        if name == "mala":
            out, acc, _ = SM.mala(post, th0, 1500, L, rng, adapt_until=300)
        else:
            out, acc, _ = SM.rwm_batched(post, th0, 3000, L, rng, adapt_until=600)
        burn = len(out) // 5
        x = out[burn:].reshape(-1, 6)
        emp = np.cov(x, rowvar=False)
        # relative error of the covariance in the metric of the target
        w = np.linalg.cholesky(np.linalg.inv(cov))
        err = np.abs(w.T @ emp @ w - np.eye(6)).max()
        assert err < 0.12, (name, err)
        assert np.abs(np.linalg.solve(np.linalg.cholesky(cov), x.mean(0) - mu)).max() < 0.1


def test_ess_autocorr_ar1():
    """Autocorrelation ESS against the exact AR(1) value n C (1 - phi)/(1 + phi):
    a long single chain, many chains, and a strongly correlated chain where
    batch means saturate near the number of batches (sqrt(n))."""
    rng = np.random.default_rng(8)
    for n, c, phi, tol in [(100000, 1, 0.9, 0.1), (1000, 200, 0.8, 0.1), (20000, 1, 0.99, 0.3)]:
        x = _ar1_chains(rng, n, c, phi)
        exact = n * c * (1 - phi) / (1 + phi)
        assert abs(SM.ess_autocorr(x) / exact - 1) < tol, (n, c, phi)
    x = _ar1_chains(rng, 20000, 1, 0.995)
    assert SM.ess_batch_means(x[:, 0]) > 2.5 * 20000 * 0.005 / 1.995
    assert SM.ess_autocorr(x) < 2.0 * 20000 * 0.005 / 1.995


class SurrogateGauss:
    """A deliberately wrong surrogate of Gauss(MU, SD): shifted mean and
    wider spread, with the central-difference interface of
    SurrogatePosterior.logpost_pet_grad."""

    def __init__(self, mu, sd):
        self.mu = np.asarray(mu, float) + 0.6 * SD
        self.sd = np.asarray(sd, float) * 1.4
        self.n_evals = 0

    def _lp(self, th):
        return -0.5 * np.sum(((th - self.mu) / self.sd) ** 2, 1)

    def logpost_pet_grad(self, th, step):
        th = np.atleast_2d(th)
        self.n_evals += 13 * len(th)
        g = np.zeros_like(th)
        for k in range(6):
            e = np.zeros(6)
            e[k] = step[k]
            g[:, k] = (self._lp(th + e) - self._lp(th - e)) / (2 * step[k])
        return self._lp(th), g


def test_da_mala_targets_the_full_posterior():
    """Delayed acceptance with Langevin proposals on a wrong surrogate still
    samples the full posterior (the second stage corrects it); dropping the
    second stage would sample the surrogate's shifted, wider Gaussian."""
    rng = np.random.default_rng(4)
    post = Gauss(MU, SD)
    spost = SurrogateGauss(MU, SD)
    th0 = MU + SD * rng.standard_normal((32, 6))
    out, acc, _, n_full, acc1 = SM.da_mala(post, spost, th0, 3000, SD, rng, fd_step=1e-3 * SD, adapt_until=500)
    x = out[500:].reshape(-1, 6)
    assert n_full < 3000 * 32 and 0.4 < acc[500:].mean() < 0.75 and acc1[500:].mean() > acc[500:].mean()
    assert np.all(np.abs(x.mean(0) - MU) < 0.1 * SD)
    assert np.all(np.abs(x.std(0) / SD - 1) < 0.08)


def test_surrogate_inputs_position_maps():
    """The surrogate's position map (a separable product of three periodic
    1-D Gaussian factors) against the 3-D Gaussian of the squared periodic
    distance, evaluated directly; the scratch-buffer path gives the same."""

    class NoSurrogate:
        batch = 4

        def set_observations(self, *a):
            pass

    lo, hi = np.zeros(6), np.full(6, 200.0)
    sp = SM.SurrogatePosterior(NoSurrogate(), (None, None, None), lo, hi, ratio=2, norm_mu=np.zeros(3),
                               norm_sd=np.ones(3))
    th = np.array([[0.1, 0.2, 10.0, 64.0, 127.0, 0.1], [0.3, 0.4, 1.0, 126.5, 70.25, 0.2],
                   [0.5, 0.6, 63.0, 0.0, 33.3, 0.3]])
    p, u = sp.inputs(th)
    n = 64
    i = np.arange(n, dtype=np.float64)
    for b, t in enumerate(th):
        d = [i - t[2 + a] / 2 for a in range(3)]
        d = [x - n * np.floor(x / n + 0.5) for x in d]
        dz, dy, dx = np.meshgrid(d[2], d[1], d[0], indexing="ij")
        ref = np.exp(-(dx * dx + dy * dy + dz * dz) / (2 * sp.sigma ** 2))
        assert np.allclose(u[b], ref, rtol=1e-6, atol=1e-7), b
        # unit peak at the nearest voxel of the centre (periodic)
        assert u[b].max() <= 1.0 and u[b].max() > 0.5
    assert np.allclose(p[:, :3], th[:, [0, 1, 5]])
    p2, u2 = sp._inputs_scratch(th)
    assert np.array_equal(u2, u) and np.array_equal(p2, p)
