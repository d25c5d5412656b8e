"""GPU log-likelihoods (FP64 accumulation on the device) against the oracle.

Reference model (HGG_Likelihood): the oracle accumulates in long double as
the reference; the device in FP64 -> relative 1e-10.  Flip-rate model:
FP64 on both sides, different summation order -> relative 1e-11.
"""
import numpy as np
import pytest

import gliomauq_amd as gq
from gliomauq_amd import anatomy as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PARAMS = [0.000361, 0.85, 0.7, 0.25, 0.09]


def grid_with_field(ctx, n, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*(np.arange(n) / n,) * 3, indexing="ij")
    u = np.clip(1.2 * np.exp(-((x - 0.5) ** 2 + (y - 0.45) ** 2 + (z - 0.55) ** 2) / 0.04) - 0.05, 0, 1)
    u = (u + 0.01 * rng.random(u.shape)).astype(np.float32)
    g = gq.Grid(ctx, n // 16, int(np.log2(n // 16)), 2)
    g.sample_volume("phi", u)
    return g, u.astype(np.float64)


@pytest.mark.parametrize("n", [64, 128])
def test_flip_likelihood(ctx, n):
    g, u = grid_with_field(ctx, n, 1)
    t1, t2, pet = A.observations(np.clip(u * 1.1, 0, 1), seed=3)
    lik = gq.FlipLikelihood(g, t1, t2, pet)
    out = lik.loglik()
    ref, parts = O.lik_flip(u, t1, t2, pet)
    assert abs(out[0] - ref) <= 1e-11 * abs(ref)
    np.testing.assert_allclose(out[1:], parts, rtol=1e-11)


@pytest.mark.parametrize("selective", [True, False])
def test_reference_likelihood(ctx, selective):
    n = 64
    g, u = grid_with_field(ctx, n, 2)
    rng = np.random.default_rng(5)
    pet = np.where(u > 0.2, u / PARAMS[1] + 0.01 * rng.standard_normal(u.shape), 0.0)
    t1 = (u > 0.65).astype(np.float64)
    t2 = (u > 0.3).astype(np.float64)
    pts = rng.integers(0, n, size=(5000, 3)).astype(np.int32)  # x, y, z
    lik = gq.HGG_Likelihood(g, pet, t1, t2, PARAMS, points_xyz=pts, selective=selective)
    out = lik.run()
    ref, parts = O.lik_ref(u, pet, t1, t2, PARAMS, selective=selective, points=np.ascontiguousarray(
        pts.T.astype(np.float64)))
    assert abs(out[0] - ref) <= 1e-10 * abs(ref)
    np.testing.assert_allclose(out[1:], parts, rtol=1e-10)


def test_likelihood_needs_observations(ctx):
    g, u = grid_with_field(ctx, 32, 0)  # no observations of size 32^3 uploaded
    out = np.zeros(4)
    params = np.array([0.7, 0.25, 0.05, 0.05])
    with pytest.raises(gq._lib.GQError) as e:
        g.ctx.call("gq_loglik_flip", gq._lib.ptr(params, gq._lib.C.c_double), gq._lib.ptr(out, gq._lib.C.c_double))
    assert e.value.status == 4
