"""Benchmark: the wavelet-adaptive FK forward solve + likelihood on one MI355X.

Workload (default): the 256^3-equivalent wavelet-adaptive single solve --
synthetic anatomy (seed 0) on a 17 cm cube, Dw = 0.0013 cm^2/day,
rho = 0.025/day, dt = 0.9 * CFL of the finest level (0.5088 days), blocks of
16^3 on levels <= 4, Interp4 refine/compress with tolerances 1e-3 / 1e-4
every 10 steps, FP32 field, then the final refinement and the flip-rate
likelihood (FP64) against synthetic observations of a ground-truth run at
perturbed parameters (Dw x1.25, rho x0.8).

A bench "step" is one FK time step of that solve (with its share of the
adaptation work).  The timed job is always the complete solve to T = 300
days (590 steps): --steps is recorded but a partial solve would time only the
uniform start of the grid, so it is rounded up to the whole solve.  The W
warmup steps run on a separate copy of the grid.  `value` is leaf
cell-updates per second of wall time over the timed region (inputs resident
in HBM).  The roofline object is for the step kernel (k_fk_step): 16
algorithmic bytes per cell update (read u, Pwm, Pgm; write u) over its
event-timed device duration.  The CPU baseline is the reference's own RD
functor and UpdateTumor compiled from source (oracle/_ref, block-wise labs,
OpenMP) on the first, still uniform, steps of the same solve.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
L_CM = 17.0


def ordinal(k):
    """'' for 1, else '2nd ', '3rd ', '4th ', ... (every k-th launch)."""
    return "" if k == 1 else {2: "2nd ", 3: "3rd "}.get(k, "%dth " % k)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pmc_traffic_ratio(workload="adaptive256"):
    """HBM traffic / algorithmic bytes of k_fk_step from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE runs of this bench,
    calibrated on a copy kernel of known size); None if absent."""
    path = os.path.join(ROOT, "profiles", "k_fk_step_%s_traffic.md" % workload)
    try:
        for line in open(path):
            if line.startswith("{") and "traffic_over_algorithmic" in line:
                return float(json.loads(line)["traffic_over_algorithmic"])
    except (OSError, ValueError):
        pass
    return None


def cpu_baseline(pwm, pgm, u0, Dw, Dg, rho, dt, target_s=12.0, what="the initial uniform grid of this solve"):
    """CPU baseline on a bounded sample: the reference's own RD functor and
    UpdateTumor (compiled from its sources into oracle/_ref, Real = float,
    cells with the CCell_extended footprint, OpenMP over blocks) on uniform
    steps of the same grid, as many as fit about `target_s` seconds (plus a
    single-core sample of a quarter of the grid, about a quarter of that);
    the FP64 oracle port where oracle/_ref is absent."""
    from oracle import oracle as O
    from oracle import ref as R
    threads = min(16, os.cpu_count() or 1)
    n = u0.shape[0]
    f32 = np.float32
    if R.available("libgqref_fk_f32.so"):
        def timed(a, nn, steps, thr):
            # wall of `steps` steps with the set-up (one step's call minus its step) subtracted
            t0 = time.perf_counter()
            R.uniform_steps(a[0], a[1], a[2], 1.0 / nn, f32(Dw), f32(Dg), f32(rho), dt, steps + 1, threads=thr)
            t1 = time.perf_counter()
            R.uniform_steps(a[0], a[1], a[2], 1.0 / nn, f32(Dw), f32(Dg), f32(rho), dt, 1, threads=thr)
            t2 = time.perf_counter()
            return (t1 - t0) - (t2 - t1)

        a = [np.ascontiguousarray(v, f32) for v in (u0, pwm, pgm)]
        timed(a, n, 1, threads)  # warm
        per = timed(a, n, 4, threads) / 4
        n_steps = int(max(4, min(2000, target_s / max(per, 1e-6))))
        wall = timed(a, n, n_steps, threads)
        m = max(16, n // 2)
        s1 = [np.ascontiguousarray(v[:m, :m, :m], f32) for v in (u0, pwm, pgm)]
        per1 = timed(s1, n, 2, 1) / 2
        n1 = int(max(2, min(2000, 0.25 * target_s / max(per1, 1e-6))))
        wall1 = timed(s1, n, n1, 1)
        return {"value": n ** 3 * n_steps / wall, "unit": "cell-updates/s", "cores": threads, "kind": "reference",
                "sample": "the reference's Glioma_ReactionDiffusionOperator + UpdateTumor compiled from source "
                          "(Real=float, CCell_extended-sized cells, per-block labs, OpenMP over blocks), %d uniform "
                          "steps of %s (%d^3, %.1f s)" % (n_steps, what, n, wall),
                "single_core": {"value": m ** 3 * n1 / wall1, "cores": 1,
                                "sample": "%d steps of a %d^3 grid (%.1f s)" % (n1, m, wall1)}}
    a = [np.ascontiguousarray(v, np.float64) for v in (u0, pwm, pgm)]
    t0 = time.perf_counter()
    O.fk_uniform_steps(a[0], a[1], a[2], 16, 1.0 / n, Dw, Dg, rho, dt, 2, threads=threads)
    per = (time.perf_counter() - t0) / 2
    n_steps = int(max(2, min(2000, target_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    O.fk_uniform_steps(a[0], a[1], a[2], 16, 1.0 / n, Dw, Dg, rho, dt, n_steps, threads=threads)
    wall = time.perf_counter() - t0
    return {"value": n ** 3 * n_steps / wall, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": "FP64 CPU oracle, %d uniform steps of %s (%d^3, %.1f s)" % (n_steps, what, n, wall)}


def ensemble_workload(args):
    """Batched ensemble: 256 parameter samples (seeded), each on its own
    adaptive grid (levels 1..3, finest 128^3, Interp4, tol 1e-3/1e-4 every 10
    steps), advanced together; FP64 flip-rate log-likelihood per sample."""
    import gliomauq_amd as gq
    from gliomauq_amd import anatomy as A

    n = 128
    bpd = n // 16
    S = args.samples
    log("[bench] ensemble: anatomy %d^3, %d samples" % (n, S))
    pwm, pgm, u0_gt, cvox = A.config_inputs(n, L_CM, args.seed)
    prm, steps, ic = A.ensemble_prior(S, n, L_CM, args.seed, pwm=pwm)
    h = 1.0 / n
    ctx = gq.Context(0)
    gq.set_autotune(ctx, args.autotune, args.zt)
    gq.set_step_persistent(ctx, args.persistent)
    # observations of a ground-truth solve at perturbed parameters
    gt = gq.Ensemble(ctx, 1, bpd, 3, 4)
    gt.write_volume("pw", pwm)
    gt.write_volume("pg", pgm)
    gt.write_volume("phi", u0_gt)
    Dw0 = 0.0013 * 1.25 / L_CM ** 2
    dt0 = 0.9 * h * h / (6 * Dw0)
    gt.run(np.array([[Dw0, 0.1 * Dw0, 0.025 * 0.8, dt0]]), [gq.n_steps_until(300.0, dt0)], adapt_every=10)
    u_true = gt.sample(0).dump_uniform()
    obs = A.observations(u_true, seed=args.seed)
    del gt

    def make():
        e = gq.Ensemble(ctx, S, bpd, 3, 4)
        e.write_volume("pw", pwm)
        e.write_volume("pg", pgm)
        e.ic_gaussian(ic)
        e.set_observations(*obs)
        return e

    log("[bench] ensemble warmup")
    w = make()
    w.run(prm, np.minimum(steps, 10), adapt_every=10)
    del w
    ens = make()
    log("[bench] ensemble timed: %d sample-steps" % int(steps.sum()))
    ctx.call("gq_set_step_timing", 1 if args.kernel_timing else 0)
    hs0 = gq.host_stats(ctx)
    t0 = time.perf_counter()
    st = ens.run(prm, steps, adapt_every=10, tol_refine=1e-3, tol_compress=1e-4, graph=not args.kernel_timing)
    ll = ens.loglik_flip()
    wall = time.perf_counter() - t0
    counts = ens.block_counts()
    cells = st["cell_updates"]
    res = {
        "metric": "cell-updates/s + %HBM roofline; posterior samples/hour vs plain MCMC at fixed error",
        "value": cells / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": int(steps.max()),
        "warmup": 10, "ms_per_step": wall * 1e3 / max(int(steps.max()), 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched ensemble: %d samples, each on its own adaptive grid (levels 1-3, finest "
                               "128^3, Interp4 tol 1e-3/1e-4 every 10 steps), Dw logU[2e-4,1.5e-2], rho "
                               "logU[2e-3,0.2], T U[30,300] d, FP32, FP64 flip log-likelihoods" % S,
                   "samples": S, "seed": args.seed, "sample_steps": int(steps.sum()),
                   "steps_min_max": [int(steps.min()), int(steps.max())],
                   "final_blocks_min_max": [int(counts.min()), int(counts.max())],
                   "loglik_mean": float(np.mean(ll[:, 0]))},
        "breakdown": {"device_chunk_ms": st["device_ms"], "adapt_s": st["adapt_s"], "chunks": st["chunks"],
                      "adaptations": st["adaptations"], "graph": bool(st["graph"]), "slab_depth": st["zt"],
                      "persistent": gq.persistent_info(ctx),
                      "host": {k: v - hs0[k] for k, v in gq.host_stats(ctx).items()}},
    }
    # the step kernel's own duration: the same run again with every step
    # launch timed by its dispatch events (no graphs), outside the timed region
    if args.kernel_timing:
        kst = st
    else:
        log("[bench] ensemble: kernel-timed pass")
        ek = make()
        ctx.call("gq_set_step_timing", 1)
        kst = ek.run(prm, steps, adapt_every=10, tol_refine=1e-3, tol_compress=1e-4, graph=False)
        ctx.call("gq_set_step_timing", 0)
        del ek
    launches = int(steps.max())  # one step launch per step over the samples still running
    achieved = 16.0 * kst["cell_updates"] / (kst["device_ms"] * 1e-3) / 1e9
    ratio = pmc_traffic_ratio("ensemble256")
    bpl = 16.0 * kst["cell_updates"] / launches
    res["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS,
                       "traffic": None if ratio is None else ratio * bpl, "traffic_over_algorithmic": ratio,
                       "traffic_source": "profiles/k_fk_step_ensemble256_traffic.md (rocprofv3 --pmc FETCH_SIZE, "
                                         "WRITE_SIZE over every step launch, calibrated on a copy kernel)",
                       "kernel": "k_fk_step", "launches": launches,
                       "kernel_avg_us": kst["device_ms"] * 1e3 / launches,
                       "algorithmic_bytes_per_launch": bpl, "cells_per_launch": kst["cell_updates"] / launches,
                       "algorithmic_bytes_per_cell": 16,
                       "timing": "every step launch timed by its dispatch events in a second, untimed run"}
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        Dw0 = 0.0013 / L_CM ** 2
        res["cpu_baseline"] = cpu_baseline(pwm, pgm, u0_gt, Dw0, 0.1 * Dw0, 0.025, 0.9 * h * h / (6 * Dw0),
                                           what="the 128^3 anatomy of the ensemble")
    print(json.dumps(res), flush=True)


def adjoint_workload(args):
    """Forward + discrete adjoint of 16 chains on the uniform 128^3 grid,
    2000 steps, checkpoints every 50 steps: one gradient of the flip-rate
    log-likelihood w.r.t. (log Dw, log rho, x0, y0, z0, A) per chain."""
    import gliomauq_amd as gq
    from gliomauq_amd import anatomy as A

    n = 128
    S = args.chains
    steps = args.adj_steps
    rng = np.random.default_rng(args.seed + 99)
    pwm, pgm, u0, c = A.config_inputs(n, L_CM, args.seed)
    h = 1.0 / n
    dt = 0.9 * h * h / (6 * 1.5e-2 / L_CM ** 2)  # stable for every Dw of the prior
    ctx = gq.Context(0)
    ens = gq.Ensemble(ctx, S, n // 16, 3, 4)
    ens.write_volume("pw", pwm)
    ens.write_volume("pg", pgm)
    truth = gq.Ensemble(ctx, 1, n // 16, 3, 4)
    truth.write_volume("pw", pwm)
    truth.write_volume("pg", pgm)
    ic_true = np.array([[c[2], c[1], c[0], 0.1, 2.0]])
    truth.ic_gaussian(ic_true)
    Dt = 0.0013 * 1.25 / L_CM ** 2
    truth.run(np.array([[Dt, 0.1 * Dt, 0.025 * 0.8, dt]]), [steps], adapt_every=0)
    obs = A.observations(truth.sample(0).dump_uniform(), seed=args.seed)
    del truth
    ens = gq.Ensemble(ctx, S, n // 16, 3, 4)
    ens.write_volume("pw", pwm)
    ens.write_volume("pg", pgm)
    ens.set_observations(*obs)
    Dw = np.exp(rng.uniform(np.log(2e-4), np.log(1.5e-2), S)) / L_CM ** 2
    rho = np.exp(rng.uniform(np.log(2e-3), np.log(0.2), S))
    prm = np.stack([Dw, 0.1 * Dw, rho, np.full(S, dt)], 1)
    ic = np.array([[c[2] + rng.uniform(-2, 2), c[1] + rng.uniform(-2, 2), c[0] + rng.uniform(-2, 2), 0.1, 2.0]
                   for _ in range(S)])
    log("[bench] adjoint warmup")
    ens.adjoint_flip(prm, 100, 50, ic=ic)
    log("[bench] adjoint timed: %d chains x %d steps" % (S, steps))
    t0 = time.perf_counter()
    reps = max(1, args.steps_adj_reps)
    for _ in range(reps):
        ll, gr = ens.adjoint_flip(prm, steps, 50, ic=ic)
    wall = (time.perf_counter() - t0) / reps
    st = ens.stats()
    cells = st["cell_updates"]
    res = {
        "metric": "cell-updates/s + %HBM roofline; posterior samples/hour vs plain MCMC at fixed error",
        "value": S / wall, "unit": "gradients/s (forward + adjoint per chain)", "n_gpus": 1, "steps": steps,
        "warmup": 1, "ms_per_step": wall * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "discrete adjoint: %d chains, uniform 128^3, %d steps, checkpoints every 50, "
                               "FP32 fields, FP64 gradients" % (S, steps), "chains": S, "seed": args.seed},
        "breakdown": {"wall_s": wall, "forward_ms": st["device_ms"], "backward_ms": st["chunks"],
                      "forward_cell_updates_per_s": cells / (st["device_ms"] * 1e-3),
                      "grad_norm_mean": float(np.mean(np.abs(gr)))},
    }
    print(json.dumps(res), flush=True)


def mcmc_workload(args):
    """Posterior samples/hour: MALA over --mcmc-chains chains (gradients
    from the discrete adjoint, all chains in one ensemble call) against plain
    random-walk Metropolis with one forward solve per proposal, on the same
    posterior of theta = (log Dw, log rho, x0, y0, z0, A) (uniform grid,
    flip-rate likelihood of synthetic observations).  Both samplers get the
    same full-covariance preconditioner (the covariance of a batched pilot
    population) and the same wall-time budget; ESS: min over parameters of
    the multi-chain autocorrelation estimate (per-chain batch means saturate
    at the number of batches when chains are correlated over more iterations
    than a batch holds, and are reported beside it)."""
    import gliomauq_amd as gq
    from gliomauq_amd import anatomy as A
    from gliomauq_amd import sampler as SM

    n, steps, C = args.mcmc_n, args.mcmc_steps, args.mcmc_chains
    rng = np.random.default_rng(args.seed + 5)
    pwm, pgm, u0, c = A.config_inputs(n, L_CM, args.seed)
    h = 1.0 / n
    dt_unit = 0.9 * h * h / (6 * 1.5e-2 / L_CM ** 2)  # stable for every Dw of the prior
    theta_true = np.array([np.log(0.0013 * 1.25), np.log(0.025 * 0.8) + 1.0, c[2] + 0.3, c[1] - 0.4, c[0] + 0.2, 0.1])
    lo = np.array([np.log(2e-4), np.log(2e-3), c[2] - 6, c[1] - 6, c[0] - 6, 0.02])
    hi = np.array([np.log(1.5e-2), np.log(0.2), c[2] + 6, c[1] + 6, c[0] + 6, 0.3])
    ctx = gq.Context(0)
    # observations from the model at theta_true
    truth = gq.Ensemble(ctx, 1, n // 16, int(np.log2(n // 16)), 4)
    truth.write_volume("pw", pwm)
    truth.write_volume("pg", pgm)
    truth.ic_gaussian(np.array([[*theta_true[2:5], theta_true[5], 2.0]]))
    Dt = np.exp(theta_true[0]) / L_CM ** 2
    truth.run(np.array([[Dt, 0.1 * Dt, np.exp(theta_true[1]), dt_unit]]), [steps], adapt_every=0)
    obs = A.observations(truth.sample(0).dump_uniform(), seed=args.seed)
    del truth
    # one context per posterior: a context holds one grid or ensemble at a time
    post_b = SM.Posterior(gq.Context(0), pwm, pgm, obs, steps, dt_unit, lo, hi, batch=C)
    post_1 = SM.Posterior(gq.Context(0), pwm, pgm, obs, steps, dt_unit, lo, hi, batch=1)
    # preconditioner: per parameter, the offset from the truth where the
    # log-posterior has dropped by 2 (geometric probe, batched)
    log("[bench] mcmc scale probe")
    scale = SM.probe_scales(post_b, theta_true, drop=2.0)
    log("[bench] scales %s" % scale)
    th0 = theta_true + 0.5 * scale * rng.standard_normal((C, 6))
    log("[bench] mcmc pilot")
    pil, _, _ = SM.mala(post_b, th0, args.mcmc_pilot, scale, rng, adapt_until=args.mcmc_pilot)
    # population preconditioner: covariance of a batched-RWM pilot over all chains (the posterior
    # is narrow and correlated: log Dw with log rho, the IC position with the amplitude); every
    # sampler, the plain one included, uses the same Cholesky factor
    op, _, _ = SM.rwm_batched(post_b, pil[-1], 100, scale, rng, adapt_until=100)
    chol = SM.population_cholesky(op[50:])
    start = op[-1]
    log("[bench] mcmc step-size adaptation")
    oa, _, logeps = SM.mala(post_b, start, 30, chol, rng, adapt_until=30)
    start_m = oa[-1]
    ob, _, lfb = SM.rwm_batched(post_b, start, 60, chol, rng, adapt_until=60)
    start_b = ob[-1]

    def timed_loop(step, state, budget=None):
        budget = budget or args.mcmc_budget
        xs, acs = [], []
        t0 = time.perf_counter()
        t_log = t0
        while time.perf_counter() - t0 < budget:
            state = step(state, xs, acs)
            if time.perf_counter() - t_log > 30.0:  # progress from the sampler itself
                t_log = time.perf_counter()
                log("[bench]   %.0f s, %d iterations" % (t_log - t0, sum(len(x) for x in xs)))
        return np.concatenate(xs, 0), np.concatenate(acs, 0), time.perf_counter() - t0

    log("[bench] mcmc MALA %d chains, %.0f s" % (C, args.mcmc_budget))

    def step_mala(st, xs, acs):
        o, a, le = SM.mala(post_b, st[0], 10, chol, rng, logeps=st[1])  # adapted step sizes kept
        xs.append(o)
        acs.append(a)
        return (o[-1], le)
    xm, am, wall_m = timed_loop(step_mala, (start_m, logeps))

    log("[bench] mcmc batched RWM %d chains, %.0f s" % (C, args.mcmc_budget))

    def step_rwmb(st, xs, acs):
        o, a, lf = SM.rwm_batched(post_b, st[0], 10, chol, rng, logf=st[1])
        xs.append(o)
        acs.append(a)
        return (o[-1], lf)
    xb, ab, wall_b = timed_loop(step_rwmb, (start_b, lfb))

    plain_budget = args.mcmc_plain_budget or args.mcmc_budget
    log("[bench] mcmc RWM 1 chain, %.0f s" % plain_budget)
    out_r, acc_r, logf = SM.rwm(post_1, start[0], 200, chol, rng, adapt_until=200)

    def step_rwm(st, xs, acs):
        o, a, lf = SM.rwm(post_1, st[0], 50, chol, rng, logf=st[1])
        xs.append(o[:, None, :])
        acs.append(a[:, None])
        return (o[-1], lf)
    xr, ar, wall_r = timed_loop(step_rwm, (out_r[-1], logf), plain_budget)

    rates = {}
    for name, x, a, w in (("rwm_plain", xr, ar, wall_r), ("rwm_%d" % C, xb, ab, wall_b), ("mala_%d" % C, xm, am, wall_m)):
        e = min(SM.ess_autocorr(x[:, :, p]) for p in range(6))
        e_bm = min(sum(SM.ess_batch_means(x[:, k, p]) for k in range(x.shape[1])) for p in range(6))
        e_cm = min(SM.ess_chains(x[:, :, p]) for p in range(6)) if x.shape[1] > 1 else None
        _, halves = split_ess_rate(x, w, SM.ess_autocorr)
        boot = bootstrap_ess_rate(x, w, SM.ess_autocorr, rng) if x.shape[1] > 1 else None
        rates[name] = {"ess_per_hour": e / w * 3600, "ess": e, "ess_per_hour_halves": halves,
                       "ess_per_hour_chain_bootstrap_90": boot, "ess_batch_means_sum": e_bm,
                       "ess_per_hour_batch_means_sum": e_bm / w * 3600, "ess_chain_means": e_cm,
                       "chains": int(x.shape[1]),
                       "iterations": int(x.shape[0]), "wall_s": w, "accept": float(a.mean()),
                       "posterior_mean": x.reshape(-1, 6).mean(0).tolist()}
    ref_mean = np.array(rates["rwm_plain"]["posterior_mean"])
    sd_r = xr.reshape(-1, 6).std(0)
    for k in rates:
        pm = np.array(rates[k]["posterior_mean"])
        rates[k]["speedup_vs_plain"] = rates[k]["ess_per_hour"] / max(rates["rwm_plain"]["ess_per_hour"], 1e-30)
        lo_k = min(rates[k]["ess_per_hour_halves"] + ([rates[k]["ess_per_hour_chain_bootstrap_90"][0]]
                                                       if rates[k]["ess_per_hour_chain_bootstrap_90"] else []))
        hi_k = max(rates[k]["ess_per_hour_halves"] + ([rates[k]["ess_per_hour_chain_bootstrap_90"][1]]
                                                       if rates[k]["ess_per_hour_chain_bootstrap_90"] else []))
        pl = rates["rwm_plain"]["ess_per_hour_halves"]
        rates[k]["speedup_vs_plain_interval"] = [lo_k / max(max(pl), 1e-30), hi_k / max(min(pl), 1e-30)]
        rates[k]["speedup_vs_plain_batch_means_sum"] = (rates[k]["ess_per_hour_batch_means_sum"] /
                                                        max(rates["rwm_plain"]["ess_per_hour_batch_means_sum"], 1e-30))
        rates[k]["mean_dev_over_plain_sd"] = (np.abs(pm - ref_mean) / np.maximum(sd_r, 1e-30)).tolist()
        rates[k]["mean_rel_dev_vs_plain"] = float((np.abs(pm - ref_mean) / np.abs(ref_mean)).max())
    best = max(rates, key=lambda k: rates[k]["ess_per_hour"])
    res = {
        "metric": "cell-updates/s + %HBM roofline; posterior samples/hour vs plain MCMC at fixed error",
        "value": rates[best]["ess_per_hour"], "unit": "effective posterior samples/hour", "n_gpus": 1,
        "steps": steps, "warmup": args.mcmc_pilot, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "MALA and batched RWM (%d chains, one ensemble call per iteration; MALA gradients "
                               "from the discrete adjoint) vs plain RWM (one forward solve per proposal), uniform "
                               "%d^3, %d steps, flip-rate likelihood, 6 parameters, %.0f s per sampler"
                               % (C, n, steps, args.mcmc_budget), "seed": args.seed, "best": best,
                   "ess": "min over parameters of the multi-chain autocorrelation ESS (Vehtari et al. 2021, "
                          "Geyer initial monotone sequence); batch-means sums and chain-means estimates beside it"},
        "mcmc": rates,
        "theta_true": theta_true.tolist(), "probe_scale": scale.tolist(),
        "preconditioner_sd": np.sqrt(np.diag(chol @ chol.T)).tolist(),
    }
    print(json.dumps(res), flush=True)


def langevin_workload(args):
    """16 Langevin (MALA) chains over the posterior of theta = (log Dw,
    log rho, x0, y0, z0, A) on the uniform 128^3 grid, 2000 steps, gradients
    of the flip-rate log-likelihood from the checkpointed discrete adjoint
    (checkpoints every 50 steps), per-chain step sizes tuned to 0.57
    acceptance during the warmup, then --lg-iters iterations per chain.  The
    run is resumable: after every block of iterations the chain states, step
    sizes, preconditioner, samples, sampling wall time and RNG state go to
    --lg-state (npz), and a later process continues from there.  The last
    process optionally runs plain RWM (one forward solve per proposal) on the
    same posterior for --lg-plain-budget seconds and compares ESS per hour."""
    import gliomauq_amd as gq
    from gliomauq_amd import anatomy as A
    from gliomauq_amd import sampler as SM

    n, steps, C, ck = args.lg_n, args.lg_steps, args.lg_chains, 50
    pwm, pgm, u0, c = A.config_inputs(n, L_CM, args.seed)
    h = 1.0 / n
    dt_unit = 0.9 * h * h / (6 * 1.5e-2 / L_CM ** 2)  # stable for every Dw of the prior
    theta_true = np.array([np.log(0.0013 * 1.25), np.log(0.025 * 0.8) + 1.0, c[2] + 0.3, c[1] - 0.4, c[0] + 0.2, 0.1])
    lo = np.array([np.log(2e-4), np.log(2e-3), c[2] - 6, c[1] - 6, c[0] - 6, 0.02])
    hi = np.array([np.log(1.5e-2), np.log(0.2), c[2] + 6, c[1] + 6, c[0] + 6, 0.3])
    ctx = gq.Context(0)
    truth = gq.Ensemble(ctx, 1, n // 16, int(np.log2(n // 16)), 4)
    truth.write_volume("pw", pwm)
    truth.write_volume("pg", pgm)
    truth.ic_gaussian(np.array([[*theta_true[2:5], theta_true[5], 2.0]]))
    Dt = np.exp(theta_true[0]) / L_CM ** 2
    truth.run(np.array([[Dt, 0.1 * Dt, np.exp(theta_true[1]), dt_unit]]), [steps], adapt_every=0)
    obs = A.observations(truth.sample(0).dump_uniform(), seed=args.seed)
    del truth
    post = SM.Posterior(gq.Context(0), pwm, pgm, obs, steps, dt_unit, lo, hi, batch=C, ckpt_every=ck)
    state = args.lg_state
    rng = np.random.default_rng(args.seed + 7)
    if state and os.path.exists(state):
        z = np.load(state)
        th, logeps, chol = z["theta"], z["logeps"], z["chol"]
        xs, acs = z["samples"], z["accepted"]
        wall, warm_s, n_grad = float(z["wall_s"]), float(z["warmup_s"]), int(z["grad_evals"])
        rng.bit_generator.state = json.loads(str(z["rng"]))
        log("[bench] langevin: resumed from %s at %d iterations (%.0f s sampled)" % (state, len(xs), wall))
    else:
        t0 = time.perf_counter()
        log("[bench] langevin: scale probe")
        scale = SM.probe_scales(post, theta_true, drop=2.0)
        th0 = theta_true + 0.5 * scale * rng.standard_normal((C, 6))
        log("[bench] langevin: pilot (%d MALA iterations, diagonal preconditioner)" % args.lg_pilot)
        pil, _, le = SM.mala(post, th0, args.lg_pilot, scale, rng, adapt_until=args.lg_pilot)
        log("[bench] langevin: population preconditioner (100 batched RWM iterations)")
        op, _, _ = SM.rwm_batched(post, pil[-1], 100, scale, rng, adapt_until=100)
        chol = SM.population_cholesky(op[50:])
        log("[bench] langevin: step-size tuning towards 0.57 (%d iterations)" % args.lg_tune)
        ot, at, logeps = SM.mala(post, op[-1], args.lg_tune, chol, rng, adapt_until=args.lg_tune)
        log("[bench] langevin: tuned acceptance over the last half %.3f, step sizes %s"
            % (at[args.lg_tune // 2:].mean(), np.round(np.exp(logeps), 3).tolist()))
        th = ot[-1]
        xs, acs = np.zeros((0, C, 6)), np.zeros((0, C), bool)
        wall, warm_s, n_grad = 0.0, time.perf_counter() - t0, 0

    def save():
        if state:
            np.savez(state + ".tmp.npz", theta=th, logeps=logeps, chol=chol, samples=xs, accepted=acs, wall_s=wall,
                     warmup_s=warm_s, grad_evals=n_grad, rng=json.dumps(rng.bit_generator.state))
            os.replace(state + ".tmp.npz", state)

    save()
    t_call = time.perf_counter()
    blk = 50
    while len(xs) < args.lg_iters and time.perf_counter() - t_call < args.lg_budget:
        k = min(blk, args.lg_iters - len(xs))
        t0 = time.perf_counter()
        o, a, logeps = SM.mala(post, th, k, chol, rng, logeps=logeps)  # step sizes fixed after tuning
        wall += time.perf_counter() - t0
        n_grad += C * (k + 1)
        th = o[-1]
        xs = np.concatenate([xs, o], 0)
        acs = np.concatenate([acs, a], 0)
        save()
        log("[bench] langevin: %d / %d iterations, %.0f s sampled, acceptance %.3f"
            % (len(xs), args.lg_iters, wall, acs.mean()))
    done = len(xs) >= args.lg_iters
    res_m = None
    if len(xs) >= 8:
        e = min(SM.ess_autocorr(xs[:, :, p]) for p in range(6))
        full, halves = split_ess_rate(xs, wall, SM.ess_autocorr)
        boot = bootstrap_ess_rate(xs, wall, SM.ess_autocorr, rng)
        res_m = {"chains": C, "iterations": int(len(xs)), "wall_s": wall, "warmup_s": warm_s,
                 "accept": float(acs.mean()), "accept_per_chain": acs.mean(0).tolist(),
                 "step_sizes": np.exp(logeps).tolist(), "gradients": n_grad, "gradients_per_s": n_grad / wall,
                 "ess": e, "ess_per_hour": full, "ess_per_hour_halves": halves,
                 "ess_per_hour_chain_bootstrap_90": boot, "posterior_mean": xs.reshape(-1, 6).mean(0).tolist(),
                 "posterior_sd": xs.reshape(-1, 6).std(0).tolist()}
    res_p = None
    if done and args.lg_plain_budget > 0:
        log("[bench] langevin: plain RWM, 1 chain, %.0f s" % args.lg_plain_budget)
        post_1 = SM.Posterior(gq.Context(0), pwm, pgm, obs, steps, dt_unit, lo, hi, batch=1)
        out_r, _, logf = SM.rwm(post_1, th[0], 200, chol, rng, adapt_until=200)
        xr, ar = [], []
        st = out_r[-1]
        t0 = time.perf_counter()
        t_log = t0
        while time.perf_counter() - t0 < args.lg_plain_budget:
            o, a, logf = SM.rwm(post_1, st, 50, chol, rng, logf=logf)
            st = o[-1]
            xr.append(o[:, None, :])
            ar.append(a[:, None])
            if time.perf_counter() - t_log > 30.0:
                t_log = time.perf_counter()
                log("[bench] langevin: plain RWM %.0f s, %d iterations" % (t_log - t0, 50 * len(xr)))
        wr = time.perf_counter() - t0
        xr = np.concatenate(xr, 0)
        ar = np.concatenate(ar, 0)
        full_r, halves_r = split_ess_rate(xr, wr, SM.ess_autocorr)
        pm_r = xr.reshape(-1, 6).mean(0)
        sd_r = xr.reshape(-1, 6).std(0)
        res_p = {"iterations": int(len(xr)), "wall_s": wr, "accept": float(ar.mean()),
                 "ess": min(SM.ess_autocorr(xr[:, :, p]) for p in range(6)), "ess_per_hour": full_r,
                 "ess_per_hour_halves": halves_r, "posterior_mean": pm_r.tolist(), "posterior_sd": sd_r.tolist()}
        if res_m:
            pm = np.array(res_m["posterior_mean"])
            lo_k = min(res_m["ess_per_hour_halves"] + [res_m["ess_per_hour_chain_bootstrap_90"][0]])
            hi_k = max(res_m["ess_per_hour_halves"] + [res_m["ess_per_hour_chain_bootstrap_90"][1]])
            res_m["speedup_vs_plain"] = res_m["ess_per_hour"] / max(full_r, 1e-30)
            res_m["speedup_vs_plain_interval"] = [lo_k / max(max(halves_r), 1e-30), hi_k / max(min(halves_r), 1e-30)]
            res_m["mean_dev_over_plain_sd"] = (np.abs(pm - pm_r) / np.maximum(sd_r, 1e-30)).tolist()
    res = {
        "metric": "cell-updates/s + %HBM roofline; posterior samples/hour vs plain MCMC at fixed error",
        "value": res_m["ess_per_hour"] if res_m else None, "unit": "effective posterior samples/hour",
        "n_gpus": 1, "steps": steps, "warmup": args.lg_pilot + 100 + args.lg_tune, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "Langevin (MALA) sampling: %d chains, uniform %d^3, %d steps, gradients from the "
                               "discrete adjoint (checkpoints every %d steps), step sizes tuned to 0.57 acceptance, "
                               "%d iterations per chain%s" % (C, n, steps, ck, args.lg_iters,
                                                               "" if done else " (partial: resumable)"),
                   "seed": args.seed, "complete": done,
                   "ess": "min over parameters of the multi-chain autocorrelation ESS (Vehtari et al. 2021)"},
        "mala": res_m, "rwm_plain": res_p, "theta_true": theta_true.tolist(),
    }
    print(json.dumps(res), flush=True)


def bootstrap_ess_rate(x, wall, ess_fn, rng, reps=100):
    """5 and 95 % points of the ESS per hour over resamplings of the chains
    (with replacement; x: [iteration, chain, parameter])."""
    c = x.shape[1]
    vals = []
    for _ in range(reps):
        y = x[:, rng.integers(c, size=c), :]
        vals.append(min(ess_fn(y[:, :, p]) for p in range(y.shape[2])) / wall * 3600)
    return [float(np.percentile(vals, 5)), float(np.percentile(vals, 95))]


def split_ess_rate(x, wall, ess_fn):
    """ESS per hour of a run and the interval its two halves give (x:
    [iteration, chain, parameter]; the smallest ESS over parameters)."""
    def e(y):
        return min(ess_fn(y[:, :, p]) for p in range(y.shape[2]))
    n = x.shape[0]
    full = e(x) / wall * 3600
    halves = [e(x[: n // 2]) / (wall / 2) * 3600, e(x[n // 2:]) / (wall / 2) * 3600] if n >= 8 else [full, full]
    return full, [min(halves), max(halves)]


def surrogate_workload(args):
    """Learned surrogate + delayed acceptance with the adaptive solver as the
    judge.  Full model: the FP32 wavelet-adaptive solver on 128^3 (levels 1-3,
    Interp4 tol 1e-3/1e-4 every 10 steps, final refinement), --sur-steps
    steps of a fixed dt, Gaussian IC, flip-rate likelihood against
    observations on the 64^3 grid (every other voxel).  Training fields: the
    batched adaptive ensemble's final states (256 samples per call) at
    parameters drawn around the truth, downsampled to 64^3.  The surrogate is
    trained by the bf16 MFMA kernels; then 16-chain delayed acceptance with
    Langevin proposals on the surrogate's gradient (first stage) and the
    adaptive solver (second stage) is compared with 16-chain batched RWM,
    16-chain delayed-acceptance RWM and plain one-chain RWM (one adaptive
    solve per proposal) at the same wall-time budget."""
    import gliomauq_amd as gq
    from gliomauq_amd import anatomy as A
    from gliomauq_amd import sampler as SM
    from gliomauq_amd import surrogate as SU

    n, steps, C = 128, args.sur_steps, args.chains
    rng = np.random.default_rng(args.seed + 11)
    pwm, pgm, u0, c = A.config_inputs(n, L_CM, args.seed)
    h = 1.0 / n
    dt = 0.9 * h * h / (6 * 1.5e-2 / L_CM ** 2)
    tt = np.array([np.log(0.0013 * 1.25), np.log(0.025 * 0.8) + 1.0, c[2] + 0.3, c[1] - 0.4, c[0] + 0.2, 0.1])
    lo = np.array([np.log(2e-4), np.log(2e-3), c[2] - 8, c[1] - 8, c[0] - 8, 0.02])
    hi = np.array([np.log(1.5e-2), np.log(0.2), c[2] + 8, c[1] + 8, c[0] + 8, 0.3])
    half = (slice(None, None, 2),) * 3
    nogrid = tuple(np.zeros((n // 2,) * 3, t) for t in (np.uint8, np.uint8, np.float32))

    # the batched adaptive solver generates the truth and the training fields
    gen = SM.AdaptivePosterior(gq.Context(0), pwm, pgm, nogrid, steps, dt, lo, hi, batch=args.sur_gen_batch)

    def solve_fields(ths):
        out = np.zeros((len(ths), n // 2, n // 2, n // 2), np.float32)
        t_log = time.perf_counter()
        for i in range(0, len(ths), gen.batch):
            blk = ths[i:i + gen.batch]
            gen.logpost(blk)
            for s in range(len(blk)):
                out[i + s] = gen.ens.sample(s).dump_uniform()[half]
            if time.perf_counter() - t_log > 30.0:
                t_log = time.perf_counter()
                log("[bench]   %d of %d fields" % (i + len(blk), len(ths)))
        return out

    log("[bench] surrogate: truth + %d training fields (adaptive 128^3, %d steps)"
        % (0 if args.sur_load else args.sur_fields, steps))
    u_true = solve_fields(tt[None])[0]
    obs64 = A.observations(u_true, seed=args.seed)
    span = np.array([0.15, 0.15, 1.5, 1.5, 1.5, 0.02])
    mu = tt[[0, 1, 5]]
    sd = span[[0, 1, 5]]
    if args.sur_load:
        # a surrogate trained by an earlier process of this workload (--sur-save)
        z = np.load(args.sur_load)
        sur_res = json.loads(str(z["result"]))
        del gen
        post_b = SM.AdaptivePosterior(gq.Context(0), pwm, pgm, obs64, steps, dt, lo, hi, batch=C)
        post_1 = SM.AdaptivePosterior(gq.Context(0), pwm, pgm, obs64, steps, dt, lo, hi, batch=1)
        sur = SU.Surrogate(batch=64, seed=args.seed)
        sur.set_anatomy(pwm[half], pgm[half])
        sur.set_params(z["params"])
        spost = SM.SurrogatePosterior(sur, obs64, lo, hi, ratio=2, norm_mu=mu, norm_sd=sd)
        log("[bench] surrogate: loaded %s (%d training fields, held-out error %.4f)"
            % (args.sur_load, sur_res["train_fields"], sur_res["heldout_rel_l2_mean"]))
        return surrogate_samplers(args, post_b, post_1, spost, sur_res, tt, lo, hi, span, rng)
    t0 = time.perf_counter()
    ths = tt + span * rng.uniform(-1, 1, (args.sur_fields + 256, 6))
    fields = solve_fields(ths)
    t_gen = time.perf_counter() - t0
    leaves = np.concatenate(gen.leaves)
    log("[bench] generated %d fields in %.1f s (final leaves %d..%d)" % (len(fields), t_gen, leaves.min(),
                                                                         leaves.max()))
    del gen
    # the samplers' full model: the same adaptive solver, observations at 64^3
    post_b = SM.AdaptivePosterior(gq.Context(0), pwm, pgm, obs64, steps, dt, lo, hi, batch=C)
    post_1 = SM.AdaptivePosterior(gq.Context(0), pwm, pgm, obs64, steps, dt, lo, hi, batch=1)
    sur = SU.Surrogate(batch=64, seed=args.seed)
    sur.set_anatomy(pwm[half], pgm[half])
    spost = SM.SurrogatePosterior(sur, obs64, lo, hi, ratio=2, norm_mu=mu, norm_sd=sd)
    ntr = args.sur_fields
    P_all, U0_all = spost.inputs(ths)
    log("[bench] surrogate: training %d epochs on %d fields" % (args.sur_epochs, ntr))
    t1 = time.perf_counter()
    losses = []
    fwd_ms = bwd_ms = 0.0
    nsteps_tr = 0
    B = args.sur_batch
    for ep in range(args.sur_epochs):
        perm = rng.permutation(ntr)
        lr = args.sur_lr * (0.5 ** (ep // max(1, args.sur_epochs // 3)))
        for i in range(0, ntr - B + 1, B):
            j = perm[i:i + B]
            losses.append(sur.train_step(P_all[j], U0_all[j], fields[j], lr))
            tm = sur.timing()
            fwd_ms += tm["forward_ms"]
            bwd_ms += tm["backward_ms"]
            nsteps_tr += 1
        log("[bench] epoch %d loss %.4e (%.1f s, device fwd+bwd %.2f ms/step)"
            % (ep, np.mean(losses[-(ntr // B):]), time.perf_counter() - t1, (fwd_ms + bwd_ms) / max(nsteps_tr, 1)))
    t_train = time.perf_counter() - t1
    f1, f2 = SU.conv_flops(B)
    Ph, U0h = P_all[ntr:], U0_all[ntr:]
    pv = sur.predict(Ph, U0h)
    held = fields[ntr:]
    err = np.sqrt(np.mean((pv - held) ** 2, axis=(1, 2, 3))) / np.sqrt(np.mean(held ** 2, axis=(1, 2, 3)))
    t2 = time.perf_counter()
    reps = 20
    for _ in range(reps):
        sur.predict(Ph[:64], U0h[:64])
    t_pred = (time.perf_counter() - t2) / reps
    pred_ms = sur.timing()["forward_ms"]
    f64_, _ = SU.conv_flops(64)
    sur_res = {"train_fields": ntr, "epochs": args.sur_epochs,
               "field_source": "batched adaptive solver (levels 1-3 of 128^3), 64^3",
               "gen_s": t_gen, "train_s": t_train, "train_steps": nsteps_tr,
               "loss_first": losses[0], "loss_last": float(np.mean(losses[-(ntr // B):])),
               "heldout_fields": len(held), "heldout_rel_l2_mean": float(err.mean()),
               "heldout_rel_l2_max": float(err.max()),
               "train_tflops": (f2 * nsteps_tr) / ((fwd_ms + bwd_ms) * 1e-3) / 1e12,
               "forward_tflops_batch64": f64_ / (pred_ms * 1e-3) / 1e12,
               "mfma_util_forward": f64_ / (pred_ms * 1e-3) / 2.5e15,
               "mfma_util_train": (f2 * nsteps_tr) / ((fwd_ms + bwd_ms) * 1e-3) / 2.5e15,
               "predict_ms_per_batch64": pred_ms, "predict_wall_s_per_batch64": t_pred}
    log("[bench] surrogate %s" % json.dumps(sur_res))
    del fields, P_all, U0_all
    if args.sur_save:
        np.savez(args.sur_save, params=sur.get_params(), result=json.dumps(sur_res))
        log("[bench] surrogate: saved %s" % args.sur_save)
    if args.sur_train_only:
        print(json.dumps({"metric": "surrogate training only", "surrogate": sur_res}), flush=True)
        return
    return surrogate_samplers(args, post_b, post_1, spost, sur_res, tt, lo, hi, span, rng)


def surrogate_samplers(args, post_b, post_1, spost, sur_res, tt, lo, hi, span, rng):
    """The sampler comparison of the surrogate workload (see surrogate_workload)."""
    from gliomauq_amd import sampler as SM

    C, steps = args.chains, args.sur_steps
    budget = args.mcmc_budget
    log("[bench] scale probe")
    scale = SM.probe_scales(post_b, tt, drop=2.0)
    start = tt + 0.3 * scale * rng.standard_normal((C, 6))
    log("[bench] preconditioner pilot")
    op, _, _ = SM.rwm_batched(post_b, start, 40, scale, rng, adapt_until=40)
    chol = SM.population_cholesky(op[20:])
    start = op[-1]
    fd = 1e-2 * np.sqrt(np.diag(chol @ chol.T))

    def run_loop(fn, state):
        xs, acs = [], []
        t0 = time.perf_counter()
        t_log = t0
        while time.perf_counter() - t0 < budget:
            state = fn(state, xs, acs)
            if time.perf_counter() - t_log > 30.0:
                t_log = time.perf_counter()
                log("[bench]   %.0f s, %d iterations" % (t_log - t0, sum(len(x) for x in xs)))
        return np.concatenate(xs, 0), np.concatenate(acs, 0), time.perf_counter() - t0, state

    log("[bench] plain RWM (one adaptive solve per proposal)")
    o, a, lf = SM.rwm(post_1, start[0], 20, chol, rng, adapt_until=20)

    def step_rwm(st, xs, acs):
        o, a, lf = SM.rwm(post_1, st[0], 10, chol, rng, logf=st[1])
        xs.append(o[:, None, :])
        acs.append(a[:, None])
        return (o[-1], lf)
    x_r, a_r, w_r, _ = run_loop(step_rwm, (o[-1], lf))
    log("[bench] batched RWM")
    o, a, lf = SM.rwm_batched(post_b, start, 10, chol, rng, adapt_until=10)

    def step_rwmb(st, xs, acs):
        o, a, lf = SM.rwm_batched(post_b, st[0], 3, chol, rng, logf=st[1])
        xs.append(o)
        acs.append(a)
        return (o[-1], lf)
    x_b, a_b, w_b, _ = run_loop(step_rwmb, (o[-1], lf))
    log("[bench] delayed acceptance, Langevin proposals on the surrogate gradient")
    nfull = {"da_mala_16": 0, "da_rwm_16": 0}
    diag = {"da_mala_16": [], "da_rwm_16": []}
    s1 = []
    o, a, le, nf, a1 = SM.da_mala(post_b, spost, start, 10, chol, rng, fd, adapt_until=10)

    def step_dam(st, xs, acs):
        o, a, le, nf, a1 = SM.da_mala(post_b, spost, st[0], 3, chol, rng, fd, logeps=st[1], diag=diag["da_mala_16"])
        nfull["da_mala_16"] += nf
        s1.append(a1)
        xs.append(o)
        acs.append(a)
        return (o[-1], le)
    x_dm, a_dm, w_dm, _ = run_loop(step_dam, (o[-1], le))
    log("[bench] delayed acceptance, random-walk proposals")
    o, a, lf, nf = SM.da_rwm(post_b, spost, start, 10, chol, rng, adapt_until=10)

    def step_dar(st, xs, acs):
        o, a, lf, nf = SM.da_rwm(post_b, spost, st[0], 3, chol, rng, logf=st[1], diag=diag["da_rwm_16"])
        nfull["da_rwm_16"] += nf
        xs.append(o)
        acs.append(a)
        return (o[-1], lf)
    x_dr, a_dr, w_dr, _ = run_loop(step_dar, (o[-1], lf))

    rates = {}
    runs = (("rwm_plain", x_r, a_r, w_r), ("rwm_16", x_b, a_b, w_b), ("da_mala_16", x_dm, a_dm, w_dm),
            ("da_rwm_16", x_dr, a_dr, w_dr))
    for name, x, a, w in runs:
        r, iv = split_ess_rate(x, w, SM.ess_autocorr)
        rates[name] = {"ess": r * w / 3600, "ess_per_hour": r, "ess_per_hour_halves": iv,
                       "iterations": int(x.shape[0]), "chains": int(x.shape[1]), "accept": float(a.mean()),
                       "wall_s": w, "posterior_mean": x[x.shape[0] // 4:].reshape(-1, 6).mean(0).tolist()}
    rates["da_mala_16"]["stage1_accept"] = float(np.concatenate(s1).mean()) if s1 else None
    for k in ("da_mala_16", "da_rwm_16"):
        x = dict((r[0], r[1]) for r in runs)[k]
        acc_n = int(dict((r[0], r[2]) for r in runs)[k].sum())
        rates[k]["full_solves"] = nfull[k]
        rates[k]["proposals"] = int(x.shape[0] * C)
        rates[k]["stage2_rejection_rate"] = 1.0 - acc_n / max(nfull[k], 1)
        if diag[k]:
            ds = np.concatenate([d[0] for d in diag[k]])
            df = np.concatenate([d[1] for d in diag[k]])
            ok = np.isfinite(ds) & np.isfinite(df)
            if ok.sum() > 2:
                rates[k]["stage2_corr_sur_full"] = float(np.corrcoef(ds[ok], df[ok])[0, 1])
    ref_mean = np.array(rates["rwm_plain"]["posterior_mean"])
    sd_ref = np.sqrt(np.diag(chol @ chol.T))
    for k in rates:
        m = np.array(rates[k]["posterior_mean"])
        rates[k]["mean_rel_dev_vs_plain"] = float((np.abs(m - ref_mean) / np.abs(ref_mean)).max())
        rates[k]["mean_dev_vs_plain_in_sd"] = (np.abs(m - ref_mean) / sd_ref).tolist()
    res = {
        "metric": "cell-updates/s + %HBM roofline; posterior samples/hour vs plain MCMC at fixed error",
        "value": rates["da_mala_16"]["ess_per_hour"], "unit": "effective posterior samples/hour", "n_gpus": 1,
        "steps": steps, "warmup": 10, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (surrogate) / f32 (solver)", "data": "synthetic",
        "config": {"workload": "surrogate + delayed acceptance (Langevin first stage on the surrogate), full model "
                               "the adaptive 128^3 solver (levels 1-3, %d steps), observations 64^3; surrogate "
                               "64^3 encoder-decoder (16-128 channels, 3^3), bf16 MFMA, %d adaptive-solver fields, "
                               "%d epochs; %d chains, %.0f s per sampler"
                               % (steps, sur_res["train_fields"], sur_res["epochs"], C, budget),
                   "seed": args.seed, "training_span": span.tolist(),
                   "preconditioner_sd": sd_ref.tolist(), "fd_step": fd.tolist()},
        "surrogate": sur_res,
        "mcmc": rates,
        "speedup_vs_plain": {k: rates[k]["ess_per_hour"] / max(rates["rwm_plain"]["ess_per_hour"], 1e-30)
                             for k in rates},
    }
    print(json.dumps(res), flush=True)


def surrogate_kernels_workload(args):
    """The surrogate's kernels alone, on random weights and inputs (no field
    generation): --sk-mode predict runs batch-64 predictions, train runs
    batch-8 training steps; one untimed call, then --sk-reps timed ones.  A
    short, fixed workload for kernel traces and counters."""
    from gliomauq_amd import surrogate as SU
    rng = np.random.default_rng(args.seed)
    n = SU.N
    batch = 64 if args.sk_mode == "predict" else 8
    sur = SU.Surrogate(batch=batch, seed=1)
    pw = rng.random((n, n, n)).astype(np.float32)
    sur.set_anatomy(pw, (1 - pw).astype(np.float32))
    prm = rng.random((batch, 4)).astype(np.float32)
    u0 = (0.5 * rng.random((batch, n, n, n))).astype(np.float32)
    tgt = rng.random((batch, n, n, n)).astype(np.float32)

    def call():
        if args.sk_mode == "predict":
            sur.predict(prm, u0)
        else:
            sur.train_step(prm, u0, tgt, 1e-4)

    call()
    t0 = time.perf_counter()
    for _ in range(args.sk_reps):
        call()
    wall = (time.perf_counter() - t0) / max(args.sk_reps, 1)
    tm = sur.timing()
    fwd, both = SU.conv_flops(batch)
    dev_ms = tm["forward_ms"] + tm["backward_ms"]
    flops = fwd if args.sk_mode == "predict" else both
    print(json.dumps({"metric": "surrogate convolution throughput", "value": flops / (dev_ms * 1e-3) / 1e12,
                      "unit": "TFLOP/s", "n_gpus": 1, "steps": args.sk_reps, "warmup": 1,
                      "ms_per_step": dev_ms, "host_ms_per_call": wall * 1e3, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic (random weights)",
                      "config": {"workload": "surrogate %s, batch %d, 64^3" % (args.sk_mode, batch)}}), flush=True)


def tmcmc_workload(args):
    """Transitional MCMC (the reference's sampler family) over the batched
    likelihood: a population of parameter sets evaluated --tm-batch at a time
    per ensemble call, tempered from the box prior to the posterior.  Uniform
    64^3 grid, 500 steps, flip-rate likelihood against observations on a grid
    coarser by --tm-stride."""
    import gliomauq_amd as gq
    from gliomauq_amd import anatomy as A
    from gliomauq_amd import sampler as SM

    n, steps = 64, 500
    rng = np.random.default_rng(args.seed + 21)
    pwm, pgm, u0, c = A.config_inputs(n, L_CM, args.seed)
    h = 1.0 / n
    dt = 0.9 * h * h / (6 * 1.5e-2 / L_CM ** 2)
    tt = np.array([np.log(0.0013 * 1.25), np.log(0.025 * 0.8) + 1.0, c[2] + 0.3, c[1] - 0.4, c[0] + 0.2, 0.1])
    lo = np.array([np.log(2e-4), np.log(2e-3), c[2] - 4, c[1] - 4, c[0] - 4, 0.05])
    hi = np.array([np.log(1.5e-2), np.log(0.2), c[2] + 4, c[1] + 4, c[0] + 4, 0.2])
    ctx = gq.Context(0)
    truth = gq.Ensemble(ctx, 1, n // 16, 2, 4)
    truth.write_volume("pw", pwm)
    truth.write_volume("pg", pgm)
    truth.ic_gaussian(np.array([[tt[2], tt[3], tt[4], tt[5], 2.0]]))
    Dt = np.exp(tt[0]) / L_CM ** 2
    truth.run(np.array([[Dt, 0.1 * Dt, np.exp(tt[1]), dt]]), [steps], adapt_every=0)
    s_ = args.tm_stride
    obs = tuple(o[::s_, ::s_, ::s_].copy() for o in A.observations(truth.sample(0).dump_uniform(), seed=args.seed))
    del truth
    post = SM.Posterior(gq.Context(0), pwm, pgm, obs, steps, dt, lo, hi, batch=args.tm_batch)

    def prior(m, r):
        return lo + (hi - lo) * r.random((m, 6))

    log("[bench] tmcmc population %d, batch %d" % (args.tm_pop, args.tm_batch))
    t0 = time.perf_counter()
    th, ll, stages = SM.tmcmc(post, prior, args.tm_pop, rng, cov_scale=0.5, moves=args.tm_moves,
                              max_stages=args.tm_stages)
    wall = time.perf_counter() - t0
    res = {
        "metric": "cell-updates/s + %HBM roofline; posterior samples/hour vs plain MCMC at fixed error",
        "value": post.n_evals / wall, "unit": "likelihood evaluations/s", "n_gpus": 1, "steps": steps,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "TMCMC, population %d, %d Metropolis moves per stage, uniform 64^3, 500 steps, "
                               "observations 1/%d per axis" % (args.tm_pop, args.tm_moves, s_), "seed": args.seed},
        "tmcmc": {"stages": len(stages), "final_q": stages[-1]["q"] if stages else 0.0,
                  "accept": [round(st["accept"], 3) for st in stages], "evaluations": post.n_evals, "wall_s": wall,
                  "posterior_mean": th.mean(0).tolist(), "posterior_sd": th.std(0).tolist(),
                  "theta_true": tt.tolist()},
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=590)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256, help="finest points per dimension")
    ap.add_argument("--adapt-every", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", default="adaptive",
                    choices=["adaptive", "ensemble", "adjoint", "mcmc", "surrogate", "tmcmc", "langevin",
                             "surrogate_kernels"])
    ap.add_argument("--sk-mode", default="predict", choices=["predict", "train"],
                    help="surrogate_kernels workload: batch-64 predictions or batch-8 training steps")
    ap.add_argument("--sk-reps", type=int, default=5)
    ap.add_argument("--tm-pop", type=int, default=512)
    ap.add_argument("--tm-batch", type=int, default=64)
    ap.add_argument("--tm-moves", type=int, default=3)
    ap.add_argument("--tm-stages", type=int, default=40)
    ap.add_argument("--tm-stride", type=int, default=4)
    ap.add_argument("--sur-steps", type=int, default=2000)
    ap.add_argument("--sur-fields", type=int, default=8192)
    ap.add_argument("--sur-epochs", type=int, default=6)
    ap.add_argument("--sur-batch", type=int, default=8)
    ap.add_argument("--sur-lr", type=float, default=3e-4)
    ap.add_argument("--sur-gen-batch", type=int, default=256)
    ap.add_argument("--sur-train-only", action="store_true")
    ap.add_argument("--sur-save", default="", help="npz: the trained surrogate's parameters and training figures")
    ap.add_argument("--sur-load", default="", help="npz written by --sur-save: skip generation and training")
    ap.add_argument("--sur-box", default="fixed", choices=["fixed", "probe"],
                    help="training parameters: fixed box around the mode, or mode +- K probe scales")
    ap.add_argument("--sur-box-k", type=float, default=4.0)
    ap.add_argument("--mcmc-n", type=int, default=64)
    ap.add_argument("--mcmc-steps", type=int, default=500)
    ap.add_argument("--mcmc-pilot", type=int, default=40)
    ap.add_argument("--mcmc-budget", type=float, default=60.0, help="wall seconds per sampler")
    ap.add_argument("--mcmc-plain-budget", type=float, default=0.0,
                    help="wall seconds of the plain one-chain sampler (default: --mcmc-budget)")
    ap.add_argument("--mcmc-chains", type=int, default=256, help="MALA chains of the mcmc workload")
    ap.add_argument("--lg-n", type=int, default=128)
    ap.add_argument("--lg-steps", type=int, default=2000)
    ap.add_argument("--lg-chains", type=int, default=16)
    ap.add_argument("--lg-iters", type=int, default=5000, help="Langevin iterations per chain after tuning")
    ap.add_argument("--lg-pilot", type=int, default=40)
    ap.add_argument("--lg-tune", type=int, default=200)
    ap.add_argument("--lg-budget", type=float, default=120.0, help="wall seconds of sampling in this process")
    ap.add_argument("--lg-plain-budget", type=float, default=0.0)
    ap.add_argument("--lg-state", default="", help="npz file to resume from and save to")
    ap.add_argument("--chains", type=int, default=16, help="adjoint chains")
    ap.add_argument("--adj-steps", type=int, default=2000)
    ap.add_argument("--steps-adj-reps", type=int, default=2)
    ap.add_argument("--samples", type=int, default=256, help="ensemble samples")
    ap.add_argument("--kernel-timing", action="store_true", help="event-time every step kernel (no graphs)")
    ap.add_argument("--persistent", type=int, default=0, choices=[0, 1, 2],
                    help="persistent step chunks (cooperative launch per chunk): 0 never, 1 autotuner candidate, 2 always")
    ap.add_argument("--autotune", type=int, default=1,
                    help="step autotune mode (1: slab depths; 4: also the column kernel and in-step face ghosts; "
                         "2: the slab depth --zt)")
    ap.add_argument("--zt", type=int, default=16, help="slab depth of --autotune 2/3 (16, 8, 4, 2, 1)")
    ap.add_argument("--timing", default="inline", choices=["separate", "inline"],
                    help="step-kernel durations from a second solve (separate) or from the timed solve (inline: "
                         "per-launch events, no graphs)")
    ap.add_argument("--step-graphs", type=int, default=1, help="gq_fk_steps replays graphs (1) or launches (0)")
    ap.add_argument("--timing-stride", type=int, default=3,
                    help="inline timing: events on every k-th step launch (counted across the solve; 1: all)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (samples would shard across GPUs; "
              "run --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        return ensemble_workload(args)
    if args.workload == "adjoint":
        return adjoint_workload(args)
    if args.workload == "mcmc":
        return mcmc_workload(args)
    if args.workload == "surrogate":
        return surrogate_workload(args)
    if args.workload == "tmcmc":
        return tmcmc_workload(args)
    if args.workload == "surrogate_kernels":
        return surrogate_kernels_workload(args)
    if args.workload == "langevin":
        return langevin_workload(args)

    import gliomauq_amd as gq
    from gliomauq_amd import anatomy as A

    n = args.n
    bpd = n // 16
    max_level = int(round(np.log2(bpd)))
    log("[bench] anatomy %d^3 seed %d" % (n, args.seed))
    pwm, pgm, u0, cvox = A.config_inputs(n, L_CM, args.seed)
    Dw_cm, rho = 0.0013, 0.025
    tol_r, tol_c = 1e-3, 1e-4

    ctx = gq.Context(0)
    gq.set_autotune(ctx, args.autotune, args.zt)
    gq.set_step_persistent(ctx, args.persistent)

    def fresh_grid():
        g = gq.Grid(ctx, bpd, max_level, 4)
        g.sample_volume("phi", u0)
        g.sample_volume("pw", pwm)
        g.sample_volume("pg", pgm)
        return g

    def solver(g):
        return gq.Glioma_HG_UQ(g, L_CM, cfl=0.9, adaptive=True, tol_refine=tol_r, tol_compress=tol_c,
                               adapt_every=args.adapt_every)

    s0 = solver(fresh_grid())
    Dw, Dg, rho_f, dt = s0.params(Dw_cm, rho)
    T_days = 300.0
    # the timed job is the complete solve of T = 300 days: a smaller --steps
    # is rounded up to it (a partial solve would time the uniform start only)
    n_total = gq.n_steps_until(T_days, dt)

    # observations from a ground-truth solve at perturbed parameters
    log("[bench] ground truth solve")
    gt = solver(fresh_grid())
    gt.run(Dw_cm * 1.25, rho * 0.8, T_days)
    u_true = gt.grid.dump_uniform()
    t1, t2, pet = A.observations(u_true, seed=args.seed)
    del gt

    # warmup on a separate grid
    log("[bench] warmup %d steps" % args.warmup)
    gw = fresh_grid()
    gq.fk_steps(gw, Dw, Dg, rho_f, dt, max(args.warmup, 1))
    del gw

    uniform_blocks = bpd ** 3

    def solve(per_kernel, progress):
        """The complete solve (Glioma_HG_UQ.run: adaptation after every full
        chunk of adapt_every steps, final refinement) and the likelihood;
        per_kernel: every step launch timed by its own dispatch events (no
        graphs).  Returns (wall s, stats, grid, loglik)."""
        g = fresh_grid()
        lik = gq.FlipLikelihood(g, t1, t2, pet)
        s = solver(g)
        gq.set_step_timing(g, per_kernel, args.timing_stride)
        st = {"kernel_ms": 0.0, "cells": 0, "all_cells": 0, "launches": 0, "adaptive_ms": 0.0, "adaptive_cells": 0,
              "adaptive_launches": 0, "steps_s": 0.0, "adapt_s": 0.0}
        gq.step_timing_drain(ctx)  # timings of earlier work are not this solve's
        done = 0
        t0 = time.perf_counter()
        while done < n_total:
            m = min(args.adapt_every, n_total - done)
            st["all_cells"] += m * g.n_blocks() * 4096
            ta = time.perf_counter()
            gq.fk_steps(g, Dw, Dg, rho_f, dt, m)  # returns once the launches are queued
            st["steps_s"] += time.perf_counter() - ta
            done += m
            if progress and (done % 100 == 0 or done == n_total):
                log("[bench] step %d/%d blocks %d" % (done, n_total, g.n_blocks()))
            if m == args.adapt_every:
                tc = time.perf_counter()
                s._adapt()
                st["adapt_s"] += time.perf_counter() - tc
        tc = time.perf_counter()
        gq.Science.AutomaticRefinement(g, tol_r, max_level, 1)
        ll = lik.loglik()
        wall = time.perf_counter() - t0
        st["adapt_s"] += time.perf_counter() - tc
        # the timed step launches, read after the solve (the solve never waits for their events)
        ms, cells = gq.step_timing_drain(ctx)
        ad = cells < uniform_blocks * 4096  # launches after the first compression
        st.update(kernel_ms=float(ms.sum()), cells=int(cells.sum()), launches=int(ms.size),
                  adaptive_ms=float(ms[ad].sum()), adaptive_cells=int(cells[ad].sum()),
                  adaptive_launches=int(ad.sum()))
        return wall, st, g, ll

    log("[bench] timed: %d steps (T = %g days)" % (n_total, T_days))
    gq.set_step_graphs(ctx, args.step_graphs)
    hs0 = gq.host_stats(ctx)
    wall, main, g, ll = solve(args.timing == "inline", True)
    hs1 = gq.host_stats(ctx)
    final_blocks = g.n_blocks()
    del g
    if args.timing == "inline":
        stats = main
    else:
        # the step kernel's own launch durations: the same solve again with every step launch
        # timed by its dispatch events (the timed solve replays graphs, which cannot time launches)
        log("[bench] kernel-timed solve")
        _, stats, gk, _ = solve(True, False)
        del gk
    cells = main["all_cells"]
    t_steps, t_adapt = main["steps_s"], main["adapt_s"]
    zt, fused, tune = gq.autotune_info(ctx)
    value = cells / wall
    kernel_s = stats["kernel_ms"] * 1e-3
    achieved = 16.0 * stats["cells"] / kernel_s / 1e9
    bytes_per_launch = 16.0 * stats["cells"] / max(stats["launches"], 1)
    ratio = pmc_traffic_ratio()
    ad_launch = max(stats["adaptive_launches"], 1)
    res = {
        "metric": "cell-updates/s + %HBM roofline; posterior samples/hour vs plain MCMC at fixed error",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": n_total,
        "steps_requested": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / n_total,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "wavelet-adaptive single solve to T = 300 days, 256^3-equivalent (levels <= 4, "
                               "16^3 blocks), Interp4 tol 1e-3/1e-4 every 10 steps, final refinement and FP64 "
                               "flip log-likelihood, FP32 field, %d steps of dt=%.4f d; the whole solve is timed "
                               "whatever --steps asks" % (n_total, dt),
                   "n": n, "seed": args.seed, "final_blocks": final_blocks, "loglik": float(ll[0])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (None if ratio is None else ratio * bytes_per_launch),
                     "traffic_over_algorithmic": ratio,
                     "traffic_source": "profiles/k_fk_step_adaptive256_traffic.md (rocprofv3 --pmc FETCH_SIZE, "
                                       "WRITE_SIZE over the 590 step launches of the timed solve, calibrated on "
                                       "a copy kernel)",
                     "kernel": "k_fk_step", "launches": stats["launches"],
                     "timing": ("every %sstep launch of the timed solve timed by its dispatch events "
                                "(hipExtLaunchKernel)" % ordinal(args.timing_stride)
                                if args.timing == "inline" else
                                "every step launch timed by its dispatch events in a second, identical solve "
                                "(the timed solve replays graphs)"),
                     "timed_launches_of": n_total,
                     "kernel_avg_us": stats["kernel_ms"] * 1e3 / max(stats["launches"], 1),
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "algorithmic_bytes_per_cell": 16,
                     "adaptive_launches": {"launches": stats["adaptive_launches"],
                                           "avg_us": stats["adaptive_ms"] * 1e3 / ad_launch,
                                           "cells_per_launch": stats["adaptive_cells"] / ad_launch,
                                           "frac": (16.0 * stats["adaptive_cells"] / max(stats["adaptive_ms"], 1e-9)
                                                    / 1e6 / HBM_PEAK_GBS)}},
    }
    res["breakdown"] = {"fk_steps_s": t_steps, "adapt_s": t_adapt, "step_kernel_s": kernel_s,
                        "host": {k: hs1[k] - hs0[k] for k in hs1}, "slab_depth": zt,
                        "face_ghosts_in_step": fused, "autotune_us": tune,
                        "persistent": gq.persistent_info(ctx), "graphs": list(gq.graph_info(ctx)),
                        "timed_solve_kernel_ms": main["kernel_ms"]}
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        res["cpu_baseline"] = cpu_baseline(pwm, pgm, u0, Dw, Dg, rho_f, dt)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
