"""Samplers over the GPU forward/adjoint path.

The reference samples with an external TMCMC code that calls the `brain` and
`likelihood` programs once per proposal (makefile/Example/runHGG.sh).  Here
the posterior over theta = (log Dw [cm^2/day], log rho [1/day], x0, y0, z0 [voxels],
A) of the Gaussian-IC Fisher-Kolmogorov model with the flip-rate likelihood is
evaluated for many parameter sets per GPU call:

  Posterior          log-likelihood (+ gradient from the discrete adjoint) of
                     a batch of parameter sets on a uniform grid, box prior
  rwm                plain random-walk Metropolis, one forward solve per
                     proposal (the reference's pattern, on the GPU)
  AdaptivePosterior  the same with the FP32 wavelet-adaptive solver
  mala               Metropolis-adjusted Langevin, many chains per call,
                     gradient from gq_adjoint_flip, per-chain step sizes
                     adapted towards an acceptance of 0.57 during burn-in
  da_rwm, da_mala    delayed acceptance with the learned surrogate as the
                     first stage (random-walk or Langevin proposals built on
                     the surrogate's gradient), the full solver second
  tmcmc              transitional MCMC (tempered stages, resampling and
                     per-stage Metropolis moves on the whole population)
  ess_autocorr       effective sample size of one or many chains from their
                     autocorrelations (the estimator the benchmarks report)
  ess_batch_means    effective sample size by batch means
  ess_chains         effective sample size from the spread of chain means
"""
import math
import time

import numpy as np

from .solver import BS, Ensemble

PARAM_NAMES = ("logDw", "logrho", "x0", "y0", "z0", "A")


class Posterior:
    """Log-posterior of the Gaussian-IC FK model on an n^3 uniform grid of a
    cube of edge L_cm with n_steps steps of dt days; box prior `lo`/`hi`."""

    def __init__(self, ctx, pwm, pgm, obs, n_steps, dt_days, lo, hi, batch, L_cm=17.0, sigma_ic=2.0,
                 ckpt_every=50, lik=(0.7, 0.25, 0.05, 0.05)):
        n = pwm.shape[0]
        self.n = n
        self.L = L_cm
        self.n_steps = int(n_steps)
        self.dt = float(dt_days)
        self.lo = np.asarray(lo, np.float64)
        self.hi = np.asarray(hi, np.float64)
        self.sigma_ic = sigma_ic
        self.ckpt = ckpt_every
        self.lik = lik
        self.batch = batch
        lvl = int(round(math.log2(n // BS)))
        self.ens = Ensemble(ctx, batch, n // BS, lvl, 4)
        self.ens.write_volume("pw", pwm)
        self.ens.write_volume("pg", pgm)
        self.ens.set_observations(*obs)
        self.n_evals = 0
        self.n_grad_evals = 0

    def _inputs(self, th):
        th = np.asarray(th, np.float64).reshape(-1, 6)
        m = th.shape[0]
        if m > self.batch:
            raise ValueError("more parameter sets than the batch")
        inside = np.all((th >= self.lo) & (th <= self.hi), axis=1)
        thc = np.where(inside[:, None], th, 0.5 * (self.lo + self.hi))  # keep the solver in range
        full = np.tile(0.5 * (self.lo + self.hi), (self.batch, 1))
        full[:m] = thc
        Dw = np.exp(full[:, 0]) / self.L ** 2
        prm = np.stack([Dw, 0.1 * Dw, np.exp(full[:, 1]), np.full(self.batch, self.dt)], 1)
        ic = np.stack([full[:, 2], full[:, 3], full[:, 4], full[:, 5], np.full(self.batch, self.sigma_ic)], 1)
        return m, inside, prm, ic

    def logpost_grad(self, th):
        m, inside, prm, ic = self._inputs(th)
        ll, gr = self.ens.adjoint_flip(prm, self.n_steps, self.ckpt, ic=ic, th1=self.lik[0], th2=self.lik[1],
                                       flip=self.lik[2], sigma=self.lik[3])
        self.n_evals += m
        self.n_grad_evals += m
        lp = np.where(inside, ll[:m, 0], -np.inf)
        return lp, gr[:m]

    def logpost(self, th):
        m, inside, prm, ic = self._inputs(th)
        self.ens.ic_gaussian(ic)
        steps = np.zeros(self.batch, np.int32)
        steps[:m] = self.n_steps  # padding slots take no steps
        self.ens.run(prm, steps, adapt_every=0)
        ll = self.ens.loglik_flip(th1=self.lik[0], th2=self.lik[1], flip=self.lik[2], sigma=self.lik[3])
        self.n_evals += m
        return np.where(inside, ll[:m, 0], -np.inf)


class AdaptivePosterior(Posterior):
    """Log-posterior of the same model with the FP32 wavelet-adaptive solver
    (Glioma_HG_UQ::run per parameter set, batched): every call starts each
    sample from the uniform base-level grid (gq_ensemble_reset), runs
    n_steps with adaptation (refine one loop, compress) after every
    `adapt_every` steps and the final refinement, and evaluates the flip-rate
    likelihood on the uniform dump of the adapted grid."""

    def __init__(self, ctx, pwm, pgm, obs, n_steps, dt_days, lo, hi, batch, L_cm=17.0, sigma_ic=2.0,
                 lik=(0.7, 0.25, 0.05, 0.05), adapt_every=10, tol_refine=1e-3, tol_compress=1e-4):
        super().__init__(ctx, pwm, pgm, obs, n_steps, dt_days, lo, hi, batch, L_cm=L_cm, sigma_ic=sigma_ic, lik=lik)
        self.pwm = np.ascontiguousarray(pwm, np.float32)
        self.pgm = np.ascontiguousarray(pgm, np.float32)
        self.adapt_every = adapt_every
        self.tol = (tol_refine, tol_compress)
        self.leaves = []

    def logpost_grad(self, th):
        raise NotImplementedError("the adjoint covers uniform grids only")

    def logpost(self, th):
        m, inside, prm, ic = self._inputs(th)
        ens = self.ens
        ens.reset()
        ens.write_volume("pw", self.pwm)
        ens.write_volume("pg", self.pgm)
        ens.ic_gaussian(ic)
        steps = np.zeros(self.batch, np.int32)
        steps[:m] = self.n_steps
        ens.run(prm, steps, adapt_every=self.adapt_every, tol_refine=self.tol[0], tol_compress=self.tol[1])
        ll = ens.loglik_flip(th1=self.lik[0], th2=self.lik[1], flip=self.lik[2], sigma=self.lik[3])
        self.leaves.append(ens.block_counts()[:m].copy())
        self.n_evals += m
        return np.where(inside, ll[:m, 0], -np.inf)


class SurrogatePosterior:
    """Log-posterior with the surrogate's prediction in place of the solver:
    the same parameters, prior box and flip-rate likelihood, observations on
    the surrogate's 64^3 grid.  The surrogate inputs are the normalised
    (log Dw, log rho, A) and a unit-peak Gaussian map at (x0, y0, z0) on the
    64^3 grid (coordinates of the model grid divided by `ratio`)."""

    def __init__(self, sur, obs64, lo, hi, ratio, sigma_ic_model=2.0, lik=(0.7, 0.25, 0.05, 0.05),
                 norm_mu=None, norm_sd=None):
        self.sur = sur
        sur.set_observations(*obs64)
        self.lo = np.asarray(lo, np.float64)
        self.hi = np.asarray(hi, np.float64)
        self.ratio = ratio
        self.sigma = sigma_ic_model / ratio
        self.lik = lik
        self.batch = sur.batch
        self.mu = np.zeros(3) if norm_mu is None else np.asarray(norm_mu, np.float64)
        self.sd = np.ones(3) if norm_sd is None else np.asarray(norm_sd, np.float64)
        self.n_evals = 0

    def inputs(self, th, out=None):
        """Surrogate inputs of parameter sets th: (normalised parameters,
        position maps); the maps go to `out` ([>= len(th)][64][64][64]
        float32) when given."""
        th = np.asarray(th, np.float64).reshape(-1, 6)
        p = np.zeros((len(th), 4), np.float32)
        p[:, :3] = (th[:, [0, 1, 5]] - self.mu) / self.sd
        n = 64
        i = np.arange(n, dtype=np.float64)
        # position map with unit peak, periodic distances: the Gaussian is the
        # product of one 1-D factor per axis
        d = i[None, None, :] - th[:, 2:5, None] / self.ratio  # [item][axis x, y, z][i]
        d -= n * np.floor(d / n + 0.5)
        g = np.exp(-(d * d) / (2 * self.sigma ** 2)).astype(np.float32)
        u0 = np.empty((len(th), n, n, n), np.float32) if out is None else out[:len(th)]
        np.multiply(g[:, 2, :, None, None], g[:, 1, None, :, None], out=u0)
        u0 *= g[:, 0, None, None, :]
        return p, u0

    def _inputs_scratch(self, th):
        # inputs of a call consumed at once, in a buffer kept between calls
        # (large fresh arrays cost more in page faults than in arithmetic)
        m = len(np.asarray(th).reshape(-1, 6))
        if getattr(self, "_buf", None) is None or len(self._buf) < m:
            self._buf = np.empty((m, 64, 64, 64), np.float32)
        return self.inputs(th, out=self._buf)

    def logpost(self, th):
        th = np.asarray(th, np.float64).reshape(-1, 6)
        inside = np.all((th >= self.lo) & (th <= self.hi), axis=1)
        p, u0 = self._inputs_scratch(th)
        ll = self.sur.loglik_flip(p, u0, self.lik)
        self.n_evals += len(th)
        return np.where(inside, ll[:, 0], -np.inf)

    def logpost_pet_grad(self, th, step):
        """Log-posterior and the gradient of its Gaussian PET term (the
        Bernoulli terms are piecewise constant in the field, as for the
        adjoint) by central differences of the surrogate's own predictions,
        all 13 parameter sets per chain in batched calls; `step` = difference
        step per parameter.  The gradient is a deterministic function of
        theta, so Langevin proposals built on it keep exact acceptance ratios."""
        th = np.asarray(th, np.float64).reshape(-1, 6)
        C = len(th)
        step = np.asarray(step, np.float64)
        rows = [th]
        for k in range(6):
            for sgn in (1.0, -1.0):
                t = th.copy()
                t[:, k] += sgn * step[k]
                rows.append(t)
        allp = np.concatenate(rows)
        p, u0 = self._inputs_scratch(allp)
        ll = self.sur.loglik_flip(p, u0, self.lik)
        self.n_evals += len(allp)
        inside = np.all((th >= self.lo) & (th <= self.hi), axis=1)
        lp = np.where(inside, ll[:C, 0], -np.inf)
        pet = ll[:, 3].reshape(13, C)
        g = np.zeros((C, 6))
        for k in range(6):
            g[:, k] = (pet[1 + 2 * k] - pet[2 + 2 * k]) / (2.0 * step[k])
        return lp, g


def _precond(scale):
    """Preconditioner helpers for a vector of standard deviations (diagonal)
    or a lower-triangular Cholesky factor L of a covariance M = L L^T:
    (noise z -> L z, gradient g -> M g, quadratic form v -> v^T M^-1 v),
    each applied along the last axis."""
    s = np.asarray(scale, np.float64)
    if s.ndim == 1:
        m = s * s
        return (lambda z: z * s), (lambda g: g * m), (lambda v: np.sum(v * v / m, -1))
    mm = s @ s.T
    li = np.linalg.inv(s)
    return (lambda z: z @ s.T), (lambda g: g @ mm), (lambda v: np.sum((v @ li.T) ** 2, -1))


def population_cholesky(samples, ridge=1e-12):
    """Cholesky factor of the covariance of pooled samples (n, 6) (a small
    relative ridge keeps it positive definite)."""
    x = np.asarray(samples, np.float64).reshape(-1, 6)
    c = np.cov(x, rowvar=False)
    c += ridge * np.diag(np.diag(c)) + 1e-300 * np.eye(6)
    return np.linalg.cholesky(c)


def da_rwm(post, spost, theta0, n_iter, scale, rng, adapt_until=0, target=0.234, logf=None, callback=None,
           diag=None):
    """Delayed-acceptance random-walk Metropolis for a population of chains:
    stage 1 accepts a symmetric proposal with min(1, pi~(t')/pi~(t)) on the
    surrogate posterior `spost`; only the survivors are solved with the full
    model `post` (one batched call) and accepted with
    min(1, pi(t') pi~(t) / (pi(t) pi~(t'))) (Christen & Fox 2005), so the
    chains target pi exactly.  Returns (samples, accepted, final log f,
    number of full solves).  With a list `diag`, the (surrogate, full)
    log-posterior differences of every proposal solved in full are appended
    (arrays of the stage-2 proposals of each iteration)."""
    th = np.asarray(theta0, np.float64).copy()
    Cn = th.shape[0]
    lp = _batched(post.logpost, th, post.batch)
    ls = spost.logpost(th)
    logf = np.zeros(Cn) if logf is None else np.array(logf, np.float64)
    out = np.zeros((n_iter, Cn, 6))
    acc = np.zeros((n_iter, Cn), bool)
    n_full = 0
    noise = _precond(scale)[0]
    for it in range(n_iter):
        prop = th + np.exp(logf)[:, None] * noise(rng.standard_normal((Cn, 6)))
        lsp = spost.logpost(prop)
        a1 = np.log(rng.random(Cn)) < np.where(np.isfinite(lsp), lsp - ls, -np.inf)
        idx = np.nonzero(a1)[0]
        a = np.zeros(Cn, bool)
        if len(idx):
            lpp = _batched(post.logpost, prop[idx], post.batch)
            n_full += len(idx)
            if diag is not None:
                diag.append((lsp[idx] - ls[idx], lpp - lp[idx]))
            r2 = (lpp - lp[idx]) - (lsp[idx] - ls[idx])
            ok = np.log(rng.random(len(idx))) < np.where(np.isfinite(r2), r2, -np.inf)
            j = idx[ok]
            a[j] = True
            th[j] = prop[j]
            lp[j] = lpp[ok]
            ls[j] = lsp[j]
        acc[it] = a
        if it < adapt_until:
            logf += (a.astype(np.float64) - target) / math.sqrt(it + 1.0)
        out[it] = th
        if callback:
            callback(it, th, lp)
    return out, acc, logf, n_full


def da_mala(post, spost, theta0, n_iter, scale, rng, fd_step, adapt_until=0, target=0.57, logeps=None, diag=None):
    """Delayed acceptance with Langevin proposals for a population of chains
    (the gradient-based first stage): theta' = theta + eps^2/2 M g~(theta) +
    eps L z with g~ the surrogate's log-posterior gradient
    (SurrogatePosterior.logpost_pet_grad) and M = L L^T (or diag(scale^2)).
    Stage 1 accepts with the surrogate's Metropolis-Hastings ratio
    pi~(t') q(t|t') / (pi~(t) q(t'|t)); the survivors are solved with the
    full model `post` in one batched call and accepted with
    pi(t') pi~(t) / (pi(t) pi~(t')) (Christen & Fox 2005), so the chains
    target pi exactly.  Step sizes adapt per chain towards `target` overall
    (two-stage) acceptance for the first `adapt_until` iterations.  Returns
    (samples, accepted, final log eps, full solves, stage-1 acceptances)."""
    th = np.asarray(theta0, np.float64).copy()
    Cn = th.shape[0]
    noise, mgrad, quad = _precond(scale)
    lp = _batched(post.logpost, th, post.batch)
    ls, gs = spost.logpost_pet_grad(th, fd_step)
    logeps = np.full(Cn, math.log(0.5)) if logeps is None else np.array(logeps, np.float64)
    out = np.zeros((n_iter, Cn, 6))
    acc = np.zeros((n_iter, Cn), bool)
    acc1 = np.zeros((n_iter, Cn), bool)
    n_full = 0
    for it in range(n_iter):
        eps = np.exp(logeps)[:, None]
        mean_f = th + 0.5 * eps ** 2 * mgrad(gs)
        prop = mean_f + eps * noise(rng.standard_normal((Cn, 6)))
        lsp, gsp = spost.logpost_pet_grad(prop, fd_step)
        mean_b = prop + 0.5 * eps ** 2 * mgrad(gsp)
        lq_f = -quad(prop - mean_f) / (2 * eps[:, 0] ** 2)
        lq_b = -quad(th - mean_b) / (2 * eps[:, 0] ** 2)
        la1 = np.where(np.isfinite(lsp), lsp - ls + lq_b - lq_f, -np.inf)
        a1 = np.log(rng.random(Cn)) < la1
        idx = np.nonzero(a1)[0]
        a = np.zeros(Cn, bool)
        if len(idx):
            lpp = _batched(post.logpost, prop[idx], post.batch)
            n_full += len(idx)
            if diag is not None:
                diag.append((lsp[idx] - ls[idx], lpp - lp[idx]))
            r2 = (lpp - lp[idx]) - (lsp[idx] - ls[idx])
            ok = np.log(rng.random(len(idx))) < np.where(np.isfinite(r2), r2, -np.inf)
            j = idx[ok]
            a[j] = True
            th[j] = prop[j]
            lp[j] = lpp[ok]
            ls[j] = lsp[j]
            gs[j] = gsp[j]
        acc[it] = a
        acc1[it] = a1
        if it < adapt_until:
            logeps += (a.astype(np.float64) - target) / math.sqrt(it + 1.0)
        out[it] = th
    return out, acc, logeps, n_full, acc1


def rwm_batched(post, theta0, n_iter, scale, rng, adapt_until=0, target=0.234, logf=None):
    """Random-walk Metropolis for a population of chains, all proposals of an
    iteration in one batched solve.  Returns (samples, accepted, log f)."""
    th = np.asarray(theta0, np.float64).copy()
    Cn = th.shape[0]
    lp = _batched(post.logpost, th, post.batch)
    logf = np.zeros(Cn) if logf is None else np.array(logf, np.float64)
    out = np.zeros((n_iter, Cn, 6))
    acc = np.zeros((n_iter, Cn), bool)
    noise = _precond(scale)[0]
    for it in range(n_iter):
        prop = th + np.exp(logf)[:, None] * noise(rng.standard_normal((Cn, 6)))
        lpp = _batched(post.logpost, prop, post.batch)
        a = np.log(rng.random(Cn)) < np.where(np.isfinite(lpp), lpp - lp, -np.inf)
        th[a] = prop[a]
        lp[a] = lpp[a]
        acc[it] = a
        if it < adapt_until:
            logf += (a.astype(np.float64) - target) / math.sqrt(it + 1.0)
        out[it] = th
    return out, acc, logf


def probe_scales(post, theta, drop=2.0, lo=1e-7, hi=1.0, n=16):
    """Per parameter k, the relative offset d (geometric grid in [lo, hi] of
    the parameter's box width) where logpost(theta + d e_k) first falls `drop`
    below logpost(theta); parameter sets are evaluated in batches."""
    theta = np.asarray(theta, np.float64)
    width = post.hi - post.lo
    grid = np.geomspace(lo, hi, n)
    rows = [theta]
    for k in range(6):
        for g in grid:
            t = theta.copy()
            t[k] += g * width[k]
            rows.append(t)
    lp = _batched(post.logpost, np.array(rows), post.batch)
    base = lp[0]
    out = np.zeros(6)
    for k in range(6):
        seg = lp[1 + k * n:1 + (k + 1) * n]
        j = int(np.argmax(seg < base - drop)) if np.any(seg < base - drop) else n - 1
        out[k] = grid[j] * width[k]
    return out


def rwm(post, theta0, n_iter, scale, rng, adapt_until=0, target=0.234, callback=None, logf=0.0):
    """Random-walk Metropolis, one chain, one forward solve per proposal.
    Gaussian proposals N(0, f^2 M) with M = diag(s^2) for a vector `scale`
    or M = L L^T for a Cholesky factor L; log f starts at `logf` and adapts
    towards `target` acceptance for the first `adapt_until` iterations.
    Returns (samples, accepted, final log f)."""
    th = np.asarray(theta0, np.float64).copy()
    lp = post.logpost(th[None])[0]
    out = np.zeros((n_iter, 6))
    acc = np.zeros(n_iter, bool)
    noise = _precond(scale)[0]
    for it in range(n_iter):
        prop = th + np.exp(logf) * noise(rng.standard_normal(6))
        lpp = post.logpost(prop[None])[0]
        if np.log(rng.random()) < lpp - lp:
            th, lp = prop, lpp
            acc[it] = True
        if it < adapt_until:
            logf += (float(acc[it]) - target) / math.sqrt(it + 1.0)
        out[it] = th
        if callback:
            callback(it, th, lp)
    return out, acc, logf


def mala(post, theta0, n_iter, scale, rng, adapt_until=0, target=0.57, callback=None, logeps=None):
    """Metropolis-adjusted Langevin for a population of chains (theta0: (C, 6)),
    preconditioner M = diag(scale^2) (vector) or L L^T (Cholesky factor L),
    per-chain step sizes exp(logeps) (default
    0.5) adapted towards `target` acceptance for the first `adapt_until`
    iterations.  Returns (samples, accepted, final logeps)."""
    th = np.asarray(theta0, np.float64).copy()
    C = th.shape[0]
    noise, mgrad, quad = _precond(scale)
    lp, g = post.logpost_grad(th)
    logeps = np.full(C, math.log(0.5)) if logeps is None else np.array(logeps, np.float64)
    out = np.zeros((n_iter, C, 6))
    acc = np.zeros((n_iter, C), bool)
    for it in range(n_iter):
        eps = np.exp(logeps)[:, None]
        mean_f = th + 0.5 * eps ** 2 * mgrad(g)
        prop = mean_f + eps * noise(rng.standard_normal((C, 6)))
        lpp, gp = post.logpost_grad(prop)
        mean_b = prop + 0.5 * eps ** 2 * mgrad(gp)
        lq_f = -quad(prop - mean_f) / (2 * eps[:, 0] ** 2)
        lq_b = -quad(th - mean_b) / (2 * eps[:, 0] ** 2)
        la = lpp - lp + lq_b - lq_f
        a = np.log(rng.random(C)) < np.where(np.isfinite(la), la, -np.inf)
        th = np.where(a[:, None], prop, th)
        lp = np.where(a, lpp, lp)
        g = np.where(a[:, None], gp, g)
        acc[it] = a
        if it < adapt_until:
            logeps += (a.astype(np.float64) - target) / math.sqrt(it + 1.0)
        out[it] = th
        if callback:
            callback(it, th, lp)
    return out, acc, logeps


def ess_batch_means(x, n_batches=None):
    """ESS of a 1-D chain by batch means (batches of ~sqrt(n) samples)."""
    x = np.asarray(x, np.float64)
    n = len(x)
    if n < 4:
        return float(n)
    b = n_batches or max(2, int(math.sqrt(n)))
    m = n // b
    xb = x[: m * b].reshape(b, m).mean(1)
    var = x.var(ddof=1)
    if var == 0:
        return 1.0  # a chain that never moved carries one sample's information
    bvar = m * xb.var(ddof=1)
    return float(min(n, n * var / bvar)) if bvar > 0 else float(n)


def ess_chains(x):
    """ESS of C parallel chains x[iteration, chain] (one parameter) from the
    spread of the chain means: C * var / var(chain means), at most the sample
    count.  Each chain is one batch, so the estimate needs no batch length
    shorter than the chains themselves; chains that have not forgotten
    dispersed starts inflate var(chain means) and lower the estimate."""
    x = np.asarray(x, np.float64)
    n, c = x.shape
    if c < 2 or n < 2:
        return float(n * c)
    var = x.var(ddof=1)
    if var == 0:
        return 1.0
    vm = x.mean(0).var(ddof=1)
    return float(min(n * c, c * var / vm)) if vm > 0 else float(n * c)


def ess_autocorr(x):
    """ESS of one or more chains x[iteration, chain] (one parameter) from the
    chains' autocorrelations: the multi-chain estimator of Vehtari, Gelman,
    Simpson, Carpenter and Buerkner (2021, without rank normalisation) --
    rho_t = 1 - (W - mean_c acov_c(t)) / var+, var+ = (n-1)/n W + B/n --
    summed with Geyer's initial positive and monotone sequence; n C / tau.
    Unlike batch means it does not saturate at the number of batches when
    the chains are correlated over more iterations than a batch holds."""
    x = np.asarray(x, np.float64)
    if x.ndim == 1:
        x = x[:, None]
    n, c = x.shape
    if n < 4:
        return float(n * c)
    xc = x - x.mean(0)
    m = 1
    while m < 2 * n:
        m *= 2
    f = np.fft.rfft(xc, m, axis=0)
    acov = np.fft.irfft(f * np.conj(f), m, axis=0)[:n] / n  # biased autocovariance per chain
    w = acov[0].mean() * n / (n - 1.0)  # mean within-chain variance
    b_n = x.mean(0).var(ddof=1) if c > 1 else 0.0
    var_plus = (n - 1.0) / n * w + b_n
    if var_plus <= 0:
        return 1.0
    rho = 1.0 - (w - acov.mean(1)) / var_plus
    rho[0] = 1.0
    tau = -1.0
    prev = np.inf
    for k in range(0, n - 1, 2):
        pk = rho[k] + rho[k + 1]
        if pk <= 0:
            break
        pk = min(pk, prev)  # monotone
        tau += 2.0 * pk
        prev = pk
    return float(min(n * c * math.log10(n * c), n * c / max(tau, 1e-12)))


def ess_parallel(x):
    """Conservative ESS of parallel chains x[iteration, chain] (one
    parameter): the smaller of the per-chain batch-means sum and the
    chain-means estimate."""
    x = np.asarray(x, np.float64)
    return min(sum(ess_batch_means(x[:, k]) for k in range(x.shape[1])), ess_chains(x))


def tmcmc(post, prior_sample, n_pop, rng, cov_scale=0.2, max_stages=50, moves=1, cov_target=1.0):
    """Transitional MCMC: tempered stages p_j ~ L^q_j prior, q_j chosen so the
    coefficient of variation of the incremental weights is cov_target;
    resampling by weight; one Metropolis move per sample per stage with the
    weighted sample covariance scaled by cov_scale^2.  The population is
    evaluated in batches of the posterior's batch size."""
    th = prior_sample(n_pop, rng)
    ll = _batched(post.logpost, th, post.batch)
    q = 0.0
    stages = []
    for _ in range(max_stages):
        lo, hi = q, 1.0
        finite = np.isfinite(ll)
        llf = np.where(finite, ll, -np.inf)
        for _ in range(60):  # bisection on the tempering increment
            mid = 0.5 * (lo + hi)
            w = np.exp((mid - q) * (llf - llf[finite].max()))
            cov = w.std() / max(w.mean(), 1e-300)
            if cov > cov_target:
                hi = mid
            else:
                lo = mid
        qn = hi if hi < 1.0 else 1.0
        w = np.exp((qn - q) * (llf - llf[finite].max()))
        w /= w.sum()
        mean = (w[:, None] * th).sum(0)
        d = th - mean
        S = (w[:, None, None] * d[:, :, None] * d[:, None, :]).sum(0) * cov_scale ** 2
        idx = rng.choice(n_pop, size=n_pop, p=w)
        th, ll = th[idx].copy(), ll[idx].copy()
        chol = np.linalg.cholesky(S + 1e-12 * np.eye(6))
        acc = 0
        for _ in range(moves):
            prop = th + rng.standard_normal((n_pop, 6)) @ chol.T
            llp = _batched(post.logpost, prop, post.batch)
            a = np.log(rng.random(n_pop)) < qn * (llp - ll)
            th[a] = prop[a]
            ll[a] = llp[a]
            acc += int(a.sum())
        stages.append({"q": qn, "accept": acc / (n_pop * moves)})
        q = qn
        if q >= 1.0:
            break
    return th, ll, stages


def _batched(fn, th, batch):
    out = np.empty(len(th))
    for i in range(0, len(th), batch):
        out[i:i + batch] = fn(th[i:i + batch])
    return out


def timed(fn, *a, **k):
    t0 = time.perf_counter()
    r = fn(*a, **k)
    return r, time.perf_counter() - t0
