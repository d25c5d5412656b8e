"""The discrete-adjoint oracle (oracle/adjoint_oracle.c) pinned by central
finite differences of the FP64 forward map (the reference's FK step restated
in oracle/fk_oracle.c, Glioma/Glioma_ReactionDiffusionOperator.h:148-266).

The reference has no adjoint.  The checks cover the parameter gradients
(log Dw with Dg = Dw/10, log rho), a random direction in u0, and the
Gaussian-IC chain rule (x0, A); one case puts CSF cells next to the tumour
so the NaN/clamp branches and the asymmetric transpose near zero-tissue
cells are exercised (a transposed coefficient or a dropped mask fails it).
"""
import numpy as np
import pytest

from gliomauq_amd import anatomy as A
from oracle import oracle as O

N = 24
L = 17.0
SIG = 0.05


def case(csf):
    pwm, pgm, _ = A.anatomy(N, L, seed=3)
    pw, pg = pwm.astype(np.float64), pgm.astype(np.float64)
    pw[:] = np.where(pw + pg > 0, pw, 0.6)
    pg[:] = np.where(pw + pg > 0, pg, 0.4)
    ic = (11.3, 12.6, 10.8, 0.5, 2.0)
    if csf:  # a CSF pocket two cells from the centre
        pw[12:14, 12:14, 14:16] = 0
        pg[12:14, 12:14, 14:16] = 0
    h = 1.0 / N
    Dw = 0.0013 / L ** 2 * 30
    rho = 0.2
    dt = 0.5 * h * h / (6 * Dw)
    rng = np.random.default_rng(0)
    pet = np.where(rng.random((N, N, N)) < 0.3, rng.random((N, N, N)) * 0.6, np.nan)
    return pw, pg, ic, h, Dw, rho, dt, pet


def loss(u, pet):
    m = ~np.isnan(pet)
    return -np.sum((pet[m] - u[m]) ** 2) / (2 * SIG ** 2)


def dloss(u, pet):
    g = np.zeros_like(u)
    m = ~np.isnan(pet)
    g[m] = (pet[m] - u[m]) / SIG ** 2
    return g


@pytest.mark.parametrize("csf", [False, True])
def test_adjoint_oracle_matches_finite_differences(csf):
    pw, pg, ic, h, Dw, rho, dt, pet = case(csf)
    steps = 30
    u0 = O.ic_gaussian(N, *ic)

    def F(Dw_, rho_, u0_):
        return loss(O.fk_uniform_steps(u0_, pw, pg, 16 if N % 16 == 0 else N, h, Dw_, 0.1 * Dw_, rho_, dt, steps),
                    pet)

    uN, d0, gD, gR = O.adjoint(u0, pw, pg, h, Dw, 0.1 * Dw, rho, dt, steps, lambda u: dloss(u, pet))
    assert abs(loss(uN, pet) - F(Dw, rho, u0)) <= 1e-9 * abs(loss(uN, pet))
    eps = 1e-4
    fd_D = (F(Dw * np.exp(eps), rho, u0) - F(Dw * np.exp(-eps), rho, u0)) / (2 * eps)
    fd_R = (F(Dw, rho * np.exp(eps), u0) - F(Dw, rho * np.exp(-eps), u0)) / (2 * eps)
    assert abs(gD - fd_D) <= 1e-5 * abs(fd_D), (gD, fd_D)
    assert abs(gR - fd_R) <= 1e-5 * abs(fd_R), (gR, fd_R)
    rng = np.random.default_rng(1)
    v = rng.standard_normal(u0.shape) * (u0 > 1e-3)
    fd_u = (F(Dw, rho, u0 + 1e-6 * v) - F(Dw, rho, u0 - 1e-6 * v)) / 2e-6
    assert abs(np.sum(d0 * v) - fd_u) <= 1e-5 * abs(fd_u)
    # Gaussian IC parameters through the chain rule
    g_ic = O.ic_grad(d0, *ic)
    for k, step in ((0, 1e-4 * ic[0]), (3, 1e-4 * ic[3])):
        p1, p2 = list(ic), list(ic)
        p1[k] += step
        p2[k] -= step
        fd = (F(Dw, rho, O.ic_gaussian(N, *p1)) - F(Dw, rho, O.ic_gaussian(N, *p2))) / (2 * step)
        assert abs(g_ic[k] - fd) <= 1e-5 * abs(fd), (k, g_ic[k], fd)


def test_adjoint_of_pure_reaction_is_closed_form():
    """Dw -> 0: u_{n+1} = u + dt rho u (1 - u) per cell, so lambda_0 is the
    product of (1 + dt rho (1 - 2 u_n)) along the trajectory."""
    n = 16
    rng = np.random.default_rng(4)
    u0 = 0.4 * rng.random((n, n, n))
    pw = np.full((n, n, n), 0.7)
    pg = np.full((n, n, n), 0.3)
    rho, dt, steps = 0.1, 0.5, 12
    ct = rng.standard_normal((n, n, n))
    uN, d0, gD, gR = O.adjoint(u0, pw, pg, 1.0 / n, 1e-30, 1e-31, rho, dt, steps, lambda u: ct)
    u = u0.copy()
    fac = np.ones_like(u)
    gr = np.zeros_like(u)
    traj = []
    for _ in range(steps):
        traj.append(u.copy())
        u = u + dt * rho * u * (1 - u)
    lam = ct.copy()
    for un in reversed(traj):
        gr += lam * dt * rho * un * (1 - un)
        lam = lam * (1 + dt * rho * (1 - 2 * un))
    assert np.allclose(uN, u, rtol=0, atol=1e-14)
    assert np.allclose(d0, lam, rtol=1e-12, atol=1e-14)
    assert abs(gR - gr.sum()) <= 1e-10 * abs(gr.sum())
    assert abs(gD) < 1e-20


def test_checkpointed_adjoint_equals_stored_states():
    """adjoint_ckpt (checkpoints + recomputation, OpenMP) against adjoint
    (every state stored): the same uN and dL/du0 to the bit, the gradient
    sums within FP64 reordering."""
    n, steps = 32, 37
    rng = np.random.default_rng(9)
    pw = rng.random((n, n, n))
    pg = 1 - pw
    pw[10:13, 4:7, 20:23] = 0
    pg[10:13, 4:7, 20:23] = 0
    u0 = O.ic_gaussian(n, 12.3, 5.2, 20.4, 0.7, 2.5)
    pet = np.where(rng.random((n, n, n)) < 0.5, rng.random((n, n, n)), np.nan)

    def ct(u):
        c = np.zeros_like(u)
        m = ~np.isnan(pet)
        c[m] = (pet[m] - u[m]) / 0.05 ** 2
        return c

    Dw = 0.02 / 17.0 ** 2
    h = 1.0 / n
    dt = 0.5 * h * h / (6 * Dw)
    a = O.adjoint(u0, pw, pg, h, Dw, 0.1 * Dw, 0.1, dt, steps, ct)
    for ck in (1, 5, 50):
        b = O.adjoint_ckpt(u0, pw, pg, h, Dw, 0.1 * Dw, 0.1, dt, steps, ct, ckpt_every=ck, threads=4)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert abs(a[2] - b[2]) <= 1e-12 * abs(a[2]) and abs(a[3] - b[3]) <= 1e-12 * abs(a[3])
    # the PET cotangent formed inside, in one forward pass
    c = O.adjoint_ckpt(u0, pw, pg, h, Dw, 0.1 * Dw, 0.1, dt, steps, ckpt_every=5, threads=4,
                       pet=pet.astype(np.float32), sigma=0.05)
    a32 = O.adjoint(u0, pw, pg, h, Dw, 0.1 * Dw, 0.1, dt, steps,
                    lambda u: np.where(np.isnan(pet), 0.0, (pet.astype(np.float32).astype(np.float64) - u) / 0.05 ** 2))
    assert np.array_equal(a32[1], c[1]) and abs(a32[2] - c[2]) <= 1e-12 * abs(a32[2])
