"""Pins of the oracle likelihoods.

* HGG_Likelihood: against the reference likelihood program compiled from
  source (oracle/_ref/likelihood), on seeded .dat inputs, and against the
  committed golden output of that program (its Likelihood.txt has 6
  decimals, the comparison tolerance);
* closed forms for both models.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref as R

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PARAMS = [0.000361, 0.85, 0.7, 0.25, 0.09]


def test_ref_model_golden():
    g = np.load(os.path.join(GOLD, "likelihood.npz"))
    pts = np.ascontiguousarray(g["points"].T)  # (3, npts): x, y, z rows
    v, _ = O.lik_ref(g["model"], g["pet"], g["t1"], g["t2"], g["params"], selective=True, points=pts)
    assert abs(-v - float(g["minus_logL_selective"])) < 1.5e-6
    v, _ = O.lik_ref(g["model"], g["pet"], g["t1"], g["t2"], g["params"], selective=False, points=pts)
    assert abs(-v - float(g["minus_logL_all"])) < 1.5e-6


@pytest.mark.skipif(not R.available("likelihood"), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [1, 2])
def test_ref_model_vs_reference_program(seed):
    rng = np.random.default_rng(seed)
    n = 12
    model = np.clip(rng.random((n, n, n)) * 1.2 - 0.1, 0, 1)
    pet = np.where(rng.random((n, n, n)) < 0.6, rng.random((n, n, n)), 0.0)
    t1 = (rng.random((n, n, n)) < 0.3).astype(np.float64)
    t2 = (rng.random((n, n, n)) < 0.5).astype(np.float64)
    pts = rng.integers(0, n, size=(150, 3)).astype(np.float64)
    for sel in (1, 0):
        ref, _ = R.run_likelihood(model, pet, t1, t2, PARAMS, points_xyz=pts.T, selective=sel, step_pet=2,
                                  step_ti=1)
        v, _ = O.lik_ref(model, pet, t1, t2, PARAMS, selective=bool(sel), step_pet=2, step_ti=1,
                         points=np.ascontiguousarray(pts.T))
        assert abs(-v - ref) < 1.5e-6


def test_ref_model_closed_form():
    """u == uc everywhere: alpha = 1/2 so each Bernoulli term is log 1/2; the
    PET term with PET*scale == u is -N/2 log(2 pi sigma^2)."""
    n = 6
    uc = PARAMS[2]
    model = np.full((n, n, n), uc)
    pet = np.full((n, n, n), uc / PARAMS[1])
    t1 = np.ones((n, n, n))
    t2 = np.zeros((n, n, n))
    v, parts = O.lik_ref(model, pet, t1, t2, [PARAMS[0], PARAMS[1], uc, 0.25, 0.09], selective=False, step_pet=3,
                         step_ti=1)
    npet = 2 ** 3
    assert abs(parts[0] - (-0.5 * npet * np.log(2 * np.pi * PARAMS[0]))) < 1e-9
    assert abs(parts[1] - n ** 3 * np.log(0.5)) < 1e-9


def test_flip_model_closed_form_and_brute_force():
    rng = np.random.default_rng(3)
    n = 8
    u = rng.random((n, n, n))
    t1 = (u > 0.7).astype(np.uint8)
    t2 = (u > 0.25).astype(np.uint8)
    pet = np.where(u > 0.25, u, np.nan).astype(np.float32)
    v, parts = O.lik_flip(u, t1, t2, pet)
    N = n ** 3
    # perfect masks: every voxel has probability 0.95
    assert abs(parts[0] - N * np.log(0.95)) < 1e-9
    assert abs(parts[1] - N * np.log(0.95)) < 1e-9
    # PET residual is only the float32 rounding of u
    npet = int((u > 0.25).sum())
    r = pet[u > 0.25].astype(np.float64) - u[u > 0.25]
    exp = npet * (-0.5 * np.log(2 * np.pi * 0.05 ** 2)) - np.sum(r * r) / (2 * 0.05 ** 2)
    assert abs(parts[2] - exp) < 1e-9 * abs(exp)
    # flipping k voxels of T1 changes the sum by k*(log 0.05 - log 0.95)
    t1b = t1.copy()
    t1b.reshape(-1)[:5] ^= 1
    v2, p2 = O.lik_flip(u, t1b, t2, pet)
    assert abs((p2[0] - parts[0]) - 5 * (np.log(0.05) - np.log(0.95))) < 1e-9


def reference_inputs_128(seed=3):
    """Seeded 128^3 volumes matching the reference's Sphere.dat coordinates
    (the points lie in [23, 114]^3): a tumour-like model field, PET on 40 % of
    the voxels, T1/T2 masks."""
    rng = np.random.default_rng(seed)
    n = 128
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    r = np.sqrt((x - 66.0) ** 2 + (y - 70.0) ** 2 + (z - 60.0) ** 2)
    model = 1.0 / (1.0 + np.exp((r - 25.0) / 4.0))
    pet = np.where(rng.random((n, n, n)) < 0.4, np.clip(model + 0.05 * rng.standard_normal((n, n, n)), 0, None), 0.)
    t1 = ((model > 0.7) ^ (rng.random((n, n, n)) < 0.05)).astype(np.float64)
    t2 = ((model > 0.25) ^ (rng.random((n, n, n)) < 0.05)).astype(np.float64)
    return model, pet, t1, t2


@pytest.mark.skipif(not R.available("likelihood"), reason="oracle/_ref not built")
def test_ref_model_with_reference_sphere_and_input():
    """The reference's own Sphere.dat (80,160 points, 10,036 on the stride-2
    lattice) and LikelihoodInput.txt: the oracle against the reference
    likelihood program compiled from source, selective and all-voxel modes."""
    fx = np.load(os.path.join(GOLD, "reference_fixtures.npz"))
    pts = fx["sphere"].astype(np.float64)  # (3, 80160)
    params = fx["likelihood_input"]
    assert pts.shape == (3, 80160) and list(params) == PARAMS
    assert int(np.sum(np.all(pts % 2 == 0, axis=0))) == 10036
    model, pet, t1, t2 = reference_inputs_128()
    for sel in (1, 0):
        want, _ = R.run_likelihood(model, pet, t1, t2, params, points_xyz=np.ascontiguousarray(pts), selective=sel)
        got, _ = O.lik_ref(model, pet, t1, t2, params, selective=bool(sel), points=np.ascontiguousarray(pts))
        print("selective=%d: reference program %.6f, oracle %.6f" % (sel, want, -got))
        assert abs(-got - want) < 1.5e-6 * max(1.0, abs(want) * 1e-6)
