"""GPU discrete adjoint (gq_adjoint_flip) of the FP32 forward solve.

Bars: (1) against the FP64 adjoint oracle (itself pinned by finite
differences in test_oracle_adjoint.py) on the same inputs: every gradient
component within 1e-5 relative to the largest component of its sample (the
device differentiates its FP32 forward map, the oracle the FP64 one);
(2) central finite differences of the device's own FP32 forward (relative
step 1e-4 in log Dw, log rho, x0, A) within 2e-3 relative (FP32 rounding of
the forward is the noise floor of the difference quotient);
(3) the log-likelihood returned with the gradient is bit-identical to a
forward-only run of the same samples.
"""
import numpy as np
import pytest

import gliomauq_amd as gq
from gliomauq_amd import anatomy as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L_CM = 17.0
SIG = 0.05


def setup(n, seed=5):
    pwm, pgm, _ = A.anatomy(n, L_CM, seed=seed)
    c = A.pick_seed_voxel(pwm, n, L_CM, seed=seed)
    truth = O.ic_gaussian(n, c[2] + 0.3, c[1] - 0.2, c[0] + 0.1, 0.6, 3.0)
    t1, t2, pet = A.observations(truth, seed=seed)
    return pwm, pgm, c, (t1, t2, pet)


def params(n, Dw_cm, rho, cfl=0.5):
    Dw = Dw_cm / L_CM ** 2
    h = 1.0 / n
    return (Dw, 0.1 * Dw, rho, cfl * h * h / (6 * Dw))


def ensemble(ctx, n, S, pwm, pgm, obs):
    ens = gq.Ensemble(ctx, S, n // 16, int(np.log2(n // 16)), 4)
    ens.write_volume("pw", pwm)
    ens.write_volume("pg", pgm)
    ens.set_observations(*obs)
    return ens


def test_adjoint_vs_fp64_oracle():
    n, steps = 32, 60
    pwm, pgm, c, obs = setup(n)
    ctx = gq.Context(0)
    S = 2
    prm = [params(n, 0.02, 0.15), params(n, 0.035, 0.08)]
    ic = [(c[2] + 0.4, c[1] + 0.1, c[0] - 0.3, 0.5, 2.0), (c[2] - 0.2, c[1] + 0.5, c[0], 0.4, 2.5)]
    ens = ensemble(ctx, n, S, pwm, pgm, obs)
    ll, gr = ens.adjoint_flip(np.array(prm), steps, ckpt_every=7, ic=np.array(ic))
    pet = obs[2].astype(np.float64)
    for s in range(S):
        u0 = O.ic_gaussian(n, *ic[s]).astype(np.float32).astype(np.float64)

        def ct(u):
            g = np.zeros_like(u)
            m = ~np.isnan(pet)
            g[m] = (pet[m] - u[m]) / SIG ** 2
            return g

        Dw, Dg, rho, dt = prm[s]
        uN, d0, gD, gR = O.adjoint(u0, pwm.astype(np.float64), pgm.astype(np.float64), 1.0 / n, Dw, Dg, rho, dt,
                                   steps, ct)
        ref = np.concatenate([[gD, gR], O.ic_grad(d0, *ic[s])])
        scale = np.max(np.abs(ref))
        err = np.abs(gr[s] - ref) / scale
        print("sample %d grad gpu %s\n          oracle %s\n   rel err %s" % (s, gr[s], ref, err))
        assert np.all(err <= 1e-5)
        # log-likelihood PET term vs the oracle's final state (FP32 vs FP64 field)
        _, parts = O.lik_flip(uN, *obs)
        assert abs(ll[s, 3] - parts[2]) <= 1e-4 * abs(parts[2])
    # the log-likelihood is the forward-only one, bit for bit
    ens2 = ensemble(ctx, n, S, pwm, pgm, obs)
    ens2.ic_gaussian(np.array(ic))
    ens2.run(np.array(prm), [steps, steps], adapt_every=0)
    assert np.array_equal(ens2.loglik_flip(), ll)
    ctx.close()


@pytest.mark.parametrize("stride", [1, 2])
def test_adjoint_vs_finite_differences(stride):
    n, steps = 64, 200
    pwm, pgm, c, obs = setup(n, seed=7)
    if stride > 1:  # observations on the grid coarser by `stride`
        obs = tuple(a[::stride, ::stride, ::stride].copy() for a in obs)
    ctx = gq.Context(0)
    ic0 = np.array([c[2] + 0.25, c[1] - 0.35, c[0] + 0.15, 0.3, 2.0])
    Dw_cm, rho = 0.05, 0.2
    ens = ensemble(ctx, n, 1, pwm, pgm, obs)
    dt = params(n, 0.1, rho)[3]  # fixed step, stable for every Dw below 0.1 cm^2/day

    def prm(Dw_cm_, rho_):
        Dw = Dw_cm_ / L_CM ** 2
        return np.array([[Dw, 0.1 * Dw, rho_, dt]])

    ll, gr = ens.adjoint_flip(prm(Dw_cm, rho), steps, ckpt_every=50, ic=ic0[None])

    def lpet(Dw_cm_, rho_, ic_):
        ens.ic_gaussian(ic_[None])
        ens.run(prm(Dw_cm_, rho_), [steps], adapt_every=0)
        return ens.loglik_flip()[0, 3]

    eps = 1e-4
    fd = [(lpet(Dw_cm * np.exp(eps), rho, ic0) - lpet(Dw_cm * np.exp(-eps), rho, ic0)) / (2 * eps),
          (lpet(Dw_cm, rho * np.exp(eps), ic0) - lpet(Dw_cm, rho * np.exp(-eps), ic0)) / (2 * eps)]
    for k in (0, 3):  # x0 and amplitude
        d = eps * ic0[k]
        p1, p2 = ic0.copy(), ic0.copy()
        p1[k] += d
        p2[k] -= d
        fd.append((lpet(Dw_cm, rho, p1) - lpet(Dw_cm, rho, p2)) / (2 * d))
    adj = np.array([gr[0, 0], gr[0, 1], gr[0, 2], gr[0, 5]])
    fd = np.array(fd)
    rel = np.abs(adj - fd) / np.abs(fd)
    print("adjoint %s\nfd      %s\nrel     %s" % (adj, fd, rel))
    assert np.all(rel <= 2e-3)
    ctx.close()


def test_adjoint_rejects_adaptive_grids_and_bad_input():
    n = 32
    pwm, pgm, c, obs = setup(n)
    ctx = gq.Context(0)
    ens = ensemble(ctx, n, 1, pwm, pgm, obs)
    p = np.array([params(n, 0.02, 0.1)])
    with pytest.raises(gq._lib.GQError) as e:
        ens.adjoint_flip(np.array([[0.0, 0.0, 0.1, 0.1]]), 5)
    assert e.value.status == 7
    with pytest.raises(gq._lib.GQError) as e:
        ens.adjoint_flip(p, 0)
    assert e.value.status == 1
    g = gq.Grid(ctx, 2, 2, 4)  # an adaptive-capable grid, refined once
    g.write_volume("pw", pwm)
    g.write_volume("pg", pgm)
    g.refine([g.getBlocksInfo()[0].block_id])
    with pytest.raises(gq._lib.GQError) as e:
        ctx.call("gq_ensemble_ic_gaussian", gq._lib.ptr(np.zeros(5), gq._lib.C.c_double))
    assert e.value.status == 6
    ctx.close()
