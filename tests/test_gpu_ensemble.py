"""GPU ensembles (gq_ensemble_*): many samples in one step launch.

Bar: every sample of an ensemble run is bit-identical (field, block count,
levels, flip log-likelihood) to the same sample run alone on a one-sample
grid with the same schedule (Glioma_HG_UQ::run, Glioma/Glioma_HG_UQ.cpp:
511-588); the one-sample path is itself held to the FP64 oracle by
test_gpu_fk / test_gpu_adaptive.  A non-adaptive ensemble is also checked
directly against the FP64 oracle (relative L2 <= 1e-5, max-abs <= 1e-5).
"""
import numpy as np
import pytest

import gliomauq_amd as gq
from gliomauq_amd import anatomy as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L_CM = 17.0
N = 128


@pytest.fixture(scope="module")
def inputs():
    pwm, pgm, u0, c = A.config_inputs(N, L_CM, 3)
    rng = np.random.default_rng(11)
    wm = np.argwhere(pwm > 0.5)
    ics = []
    for k in range(4):
        z, y, x = wm[rng.integers(len(wm))]
        ics.append(A.initial_condition(N, (x, y, z), 0.1, 2.0).astype(np.float32))
    t1, t2, pet = A.observations(np.maximum(ics[0], ics[1]) * 8.0, seed=2)
    return pwm, pgm, ics, (t1, t2, pet)


def sample_params(k):
    Dw_cm = [2.4e-3, 9e-3, 6e-4, 1.3e-3][k]
    rho = [0.05, 0.02, 0.15, 0.025][k]
    Dw = Dw_cm / L_CM ** 2
    h = 1.0 / N
    return Dw, 0.1 * Dw, rho, 0.9 * h * h / (6 * Dw)


STEPS = [23, 40, 7, 30]
K = 10
TOL_R, TOL_C = 1e-3, 1e-4


def run_single(ctx, pwm, pgm, u0, prm, n_steps, adaptive=True):
    g = gq.Grid(ctx, N // 16, 3, 4)
    g.write_volume("phi", u0)
    g.write_volume("pw", pwm)
    g.write_volume("pg", pgm)
    done = 0
    while done < n_steps:
        m = min(K if adaptive else n_steps, n_steps - done)
        gq.fk_steps(g, *prm, m)
        done += m
        if adaptive and m == K:
            gq.Science.AutomaticRefinement(g, TOL_R, 3, 1)
            gq.Science.AutomaticCompression(g, TOL_C, -1)
    if adaptive:
        gq.Science.AutomaticRefinement(g, TOL_R, 3, 1)
    return g


def advance_single(g, prm, n_steps):
    """The ensemble's per-sample schedule on one grid: adaptation after every
    full chunk of K steps, final refinement."""
    done = 0
    while done < n_steps:
        m = min(K, n_steps - done)
        gq.fk_steps(g, *prm, m)
        done += m
        if m == K:
            gq.Science.AutomaticRefinement(g, TOL_R, 3, 1)
            gq.Science.AutomaticCompression(g, TOL_C, -1)
    gq.Science.AutomaticRefinement(g, TOL_R, 3, 1)


def make_ensemble(ctx, pwm, pgm, ics):
    ens = gq.Ensemble(ctx, len(ics), N // 16, 3, 4)
    ens.write_volume("pw", pwm)
    ens.write_volume("pg", pgm)
    for s, u0 in enumerate(ics):
        ens.write_volume("phi", u0, sample=s)
    return ens


def test_ensemble_equals_single_runs(inputs):
    pwm, pgm, ics, obs = inputs
    ctx = gq.Context(0)
    singles = []
    for s, u0 in enumerate(ics):
        g = run_single(ctx, pwm, pgm, u0, sample_params(s), STEPS[s])
        lv, ix, _ = g.levels_indices()
        lik = gq.FlipLikelihood(g, *obs).loglik()
        singles.append((g.dump_uniform(), g.n_blocks(), sorted(zip(lv.tolist(), map(tuple, ix.tolist()))), lik))
    for graph in (True, False):
        ens = make_ensemble(ctx, pwm, pgm, ics)
        prm = np.array([sample_params(s) for s in range(len(ics))])
        st = ens.run(prm, STEPS, adapt_every=K, tol_refine=TOL_R, tol_compress=TOL_C, graph=graph)
        print("graph=%s stats %s" % (graph, st))
        counts = ens.block_counts()
        ll = ens.loglik_flip(*obs)
        for s in range(len(ics)):
            u_s, nb_s, lvix_s, lik_s = singles[s]
            gs = ens.sample(s)
            lv, ix, _ = gs.levels_indices()
            assert counts[s] == nb_s
            assert sorted(zip(lv.tolist(), map(tuple, ix.tolist()))) == lvix_s
            u = gs.dump_uniform()
            assert np.array_equal(u, u_s), "sample %d differs (max %g)" % (s, np.max(np.abs(u - u_s)))
            assert np.array_equal(ll[s], lik_s)
        assert st["cell_updates"] > 0 and st["adaptations"] > 0
    # the samples did evolve differently
    assert len({c for c in counts.tolist()}) > 1
    ctx.close()


def test_ensemble_second_run_with_more_running_samples(inputs):
    """Face-ghost rows and face 1/K are built only for the samples still
    stepping; a second run in which a sample that stopped early steps again
    (same work order, longer running prefix) must rebuild them for it."""
    pwm, pgm, ics, _ = inputs
    phases = [[30, 7], [30, 20]]
    singles = []
    for s in range(2):
        ctx = gq.Context(0)
        g = gq.Grid(ctx, N // 16, 3, 4)
        g.write_volume("phi", ics[s])
        g.write_volume("pw", pwm)
        g.write_volume("pg", pgm)
        for ph in phases:
            advance_single(g, sample_params(s), ph[s])
        singles.append((g.dump_uniform(), g.n_blocks()))
        ctx.close()
    ctx = gq.Context(0)
    ens = make_ensemble(ctx, pwm, pgm, ics[:2])
    prm = np.array([sample_params(s) for s in range(2)])
    for ph in phases:
        ens.run(prm, ph, adapt_every=K, tol_refine=TOL_R, tol_compress=TOL_C, graph=True)
    counts = ens.block_counts()
    for s in range(2):
        assert counts[s] == singles[s][1]
        assert np.array_equal(ens.sample(s).dump_uniform(), singles[s][0]), "sample %d differs" % s
    ctx.close()


def test_ensemble_fused_steps_uneven_counts(inputs):
    """Uniform grids with per-sample step counts 5, 8, 3: step pairs over an
    unchanged running prefix run fused; each sample ends in the ping-pong
    buffer its last launch wrote and is copied to the current one.  Same bits
    as one launch per step, with and without graph capture."""
    pwm, pgm, ics, _ = inputs
    steps = [5, 8, 3]
    out = []
    for fuse in (False, True):
        for graph in (False, True):
            ctx = gq.Context(0)
            gq.set_step_fusion(ctx, fuse)
            ens = make_ensemble(ctx, pwm, pgm, ics[:3])
            prm = np.array([sample_params(s) for s in range(3)])
            ens.run(prm, steps, adapt_every=0, graph=graph)
            out.append([ens.sample(s).dump_uniform() for s in range(3)])
            ctx.close()
    for o in out[1:]:
        for s in range(3):
            assert o[s].tobytes() == out[0][s].tobytes(), "sample %d differs" % s


def test_ensemble_persistent_bit_identical(inputs):
    """Adaptive ensemble with uneven step counts: chunks as persistent
    cooperative launches (the running prefix shrinks inside a launch) give
    the same fields, topologies and log-likelihoods as graph-captured
    per-step launches."""
    pwm, pgm, ics, obs = inputs
    steps = [23, 37, 12]
    out = []
    for persist in (0, 2):
        ctx = gq.Context(0)
        gq.set_step_persistent(ctx, persist)
        ens = make_ensemble(ctx, pwm, pgm, ics[:3])
        prm = np.array([sample_params(s) for s in range(3)])
        ens.run(prm, steps, adapt_every=K, tol_refine=TOL_R, tol_compress=TOL_C, graph=True)
        assert gq.persistent_info(ctx)[0] == (persist == 2)
        out.append(([ens.sample(s).dump_uniform() for s in range(3)], ens.block_counts().tolist(),
                    ens.loglik_flip(*obs)))
        ctx.close()
    (ua, ca, la), (ub, cb, lb) = out
    assert ca == cb
    for s in range(3):
        assert ua[s].tobytes() == ub[s].tobytes(), "sample %d differs" % s
    assert np.array_equal(la, lb)


def test_ensemble_uniform_vs_oracle(inputs):
    pwm, pgm, ics, _ = inputs
    ctx = gq.Context(0)
    ens = make_ensemble(ctx, pwm, pgm, ics[:2])
    prm = []
    for s in range(2):
        Dw, Dg, rho, dt = sample_params(s)
        prm.append((Dw * 30, Dg * 30, rho * 10, dt / 30))  # faster fronts, same CFL number
    steps = [12, 7]
    ens.run(np.array(prm), steps, adapt_every=0)
    for s in range(2):
        u = ens.sample(s).dump_uniform()
        ref = O.fk_uniform_steps(ics[s].astype(np.float64), pwm.astype(np.float64), pgm.astype(np.float64), 16,
                                 1.0 / N, *prm[s], steps[s])
        e2 = float(np.sqrt(np.mean((u - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
        einf = float(np.max(np.abs(u - ref)))
        print("sample %d: rel L2 %.3e Linf %.3e" % (s, e2, einf))
        assert e2 <= 1e-5 and einf <= 1e-5
    ctx.close()


def test_ensemble_graph_relaunch_bit_identical(inputs):
    """Repeated runs whose launches are unchanged (same step counts, same
    topology) relaunch the instantiated graph without a new capture, with
    new per-sample parameters and fields each time (as a sampler does): the
    results equal runs without graphs to the bit."""
    pwm, pgm, ics, _ = inputs
    steps = [5, 8, 3]
    runs = [np.array([sample_params((s + k) % 4) for s in range(3)]) for k in range(3)]
    out = []
    for graph in (False, True):
        ctx = gq.Context(0)
        ens = make_ensemble(ctx, pwm, pgm, ics[:3])
        res = []
        for k, prm in enumerate(runs):
            for s in range(3):
                ens.write_volume("phi", ics[(s + k) % 3], sample=s)
            ens.run(prm, steps, adapt_every=0, graph=graph)
            res.append([ens.sample(s).dump_uniform() for s in range(3)])
        out.append(res)
        ctx.close()
    for k in range(3):
        for s in range(3):
            assert out[0][k][s].tobytes() == out[1][k][s].tobytes(), "run %d sample %d differs" % (k, s)


def test_ensemble_graph_cache_alternating_shapes(inputs):
    """Runs alternating between two launch shapes (odd and even step counts,
    so the start buffer alternates too) launch each shape directly the first
    time, capture it the second time and relaunch it afterwards; every result
    equals the run without graphs to the bit."""
    pwm, pgm, ics, _ = inputs
    shapes = [[5, 8, 3], [4, 7, 3]]
    out = []
    for graph in (False, True):
        ctx = gq.Context(0)
        ens = make_ensemble(ctx, pwm, pgm, ics[:3])
        res = []
        for k in range(10):
            prm = np.array([sample_params((s + k) % 4) for s in range(3)])
            for s in range(3):
                ens.write_volume("phi", ics[(s + k) % 3], sample=s)
            ens.run(prm, shapes[k % 2], adapt_every=0, graph=graph)
            res.append([ens.sample(s).dump_uniform() for s in range(3)])
        if graph:
            cap, rel, kept = gq.graph_info(ctx)
            print("captures %d relaunches %d cached %d" % (cap, rel, kept))
            # two shapes x two start buffers at most; each key's first run is
            # direct, its second a capture, later ones relaunches
            direct = 10 - cap - rel
            assert cap <= 4 and kept == cap and 1 <= direct <= 4 and rel >= 10 - 2 * 4
        out.append(res)
        ctx.close()
    for k in range(10):
        for s in range(3):
            assert out[0][k][s].tobytes() == out[1][k][s].tobytes(), "run %d sample %d differs" % (k, s)


def test_ensemble_reset_equals_fresh_ensemble(inputs):
    """gq_ensemble_reset returns every sample to its uniform base-level grid
    in the pool it already holds: after an adaptive run, a reset, new fields
    and a second adaptive run, every sample equals the same second run on a
    freshly created ensemble to the bit (leaf sets, dumps), and the
    observations set before the reset still serve the likelihood."""
    pwm, pgm, ics, obs = inputs
    prm_a = np.array([sample_params(s) for s in range(3)])
    prm_b = np.array([sample_params((s + 1) % 4) for s in range(3)])
    steps_b = [STEPS[(s + 1) % 4] for s in range(3)]
    ctx = gq.Context(0)
    ens = make_ensemble(ctx, pwm, pgm, ics[:3])
    ens.set_observations(*obs)
    ens.run(prm_a, STEPS[:3], adapt_every=K, tol_refine=TOL_R, tol_compress=TOL_C)
    assert np.any(ens.block_counts() != (N // 16) ** 3)  # the first run adapted
    ens.reset()
    assert np.all(ens.block_counts() == (N // 16) ** 3)
    ens.write_volume("pw", pwm)
    ens.write_volume("pg", pgm)
    for s in range(3):
        ens.write_volume("phi", ics[(s + 1) % 4], sample=s)
    ens.run(prm_b, steps_b, adapt_every=K, tol_refine=TOL_R, tol_compress=TOL_C)
    got = [(ens.sample(s).levels_indices()[:2], ens.sample(s).dump_uniform()) for s in range(3)]
    ll = ens.loglik_flip()
    del ens
    fresh = make_ensemble(ctx, pwm, pgm, [ics[(s + 1) % 4] for s in range(3)])
    fresh.set_observations(*obs)
    fresh.run(prm_b, steps_b, adapt_every=K, tol_refine=TOL_R, tol_compress=TOL_C)
    for s in range(3):
        lv, ix = fresh.sample(s).levels_indices()[:2]
        assert np.array_equal(lv, got[s][0][0]) and np.array_equal(ix, got[s][0][1]), s
        assert fresh.sample(s).dump_uniform().tobytes() == got[s][1].tobytes(), "sample %d differs" % s
    assert np.array_equal(fresh.loglik_flip(), ll)
    ctx.close()


def test_ensemble_zero_step_samples(inputs):
    """Samples with zero steps beside a running one, and an ensemble where no
    sample steps: the idle samples keep their initial state to the bit, the
    running one equals its single run."""
    pwm, pgm, ics, _ = inputs
    ctx = gq.Context(0)
    prm = np.array([sample_params(s) for s in range(3)])
    ens = make_ensemble(ctx, pwm, pgm, ics[:3])
    u_init = [ens.sample(s).dump_uniform() for s in range(3)]
    ens.run(prm, [0, 0, 0], adapt_every=0)
    for s in range(3):
        assert ens.sample(s).dump_uniform().tobytes() == u_init[s].tobytes()
    ens.run(prm, [0, 5, 0], adapt_every=0)
    assert ens.sample(0).dump_uniform().tobytes() == u_init[0].tobytes()
    assert ens.sample(2).dump_uniform().tobytes() == u_init[2].tobytes()
    u1 = ens.sample(1).dump_uniform()
    g = run_single(ctx, pwm, pgm, ics[1], sample_params(1), 5, adaptive=False)
    assert np.array_equal(g.dump_uniform(), u1)
    assert not np.array_equal(u1, u_init[1])
    ctx.close()


def test_ensemble_rejects_bad_parameters(inputs):
    pwm, pgm, ics, _ = inputs
    ctx = gq.Context(0)
    ens = make_ensemble(ctx, pwm, pgm, ics[:2])
    prm = np.array([sample_params(0), sample_params(1)])
    prm[1, 3] = float("nan")
    with pytest.raises(gq._lib.GQError) as e:
        ens.run(prm, [3, 3], adapt_every=0)
    assert e.value.status == 7
    with pytest.raises(gq._lib.GQError) as e:
        ens.run(np.array([sample_params(0), sample_params(1)]), [3, -1], adapt_every=0)
    assert e.value.status == 1
    # the state is untouched by the rejected calls: a valid run still matches a single run
    ens.run(np.array([sample_params(0), sample_params(1)]), [3, 2], adapt_every=0)
    g = run_single(ctx, pwm, pgm, ics[1], sample_params(1), 2, adaptive=False)
    u1 = g.dump_uniform()
    ens2 = make_ensemble(ctx, pwm, pgm, ics[:2])
    ens2.run(np.array([sample_params(0), sample_params(1)]), [3, 2], adapt_every=0)
    assert np.array_equal(ens2.sample(1).dump_uniform(), u1)
    ctx.close()
