"""Pins of the CPU oracle's FK operator + update (oracle/fk_oracle.c).

* bit-exact against the reference's own Glioma_ReactionDiffusionOperator and
  UpdateTumor compiled from source (oracle/_ref), Real=float and Real=double;
* against committed golden vectors made by the same compiled reference
  (tests/golden/make_golden.py) so the pin holds where oracle/_ref is absent;
* closed forms that a dropped term, a wrong sign or a transposed neighbour
  would break.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref as R

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 1.0 / 128
DW = 0.0013 / 17.0 ** 2
DG = 0.1 * DW
RHO = 0.025


def random_lab(rng, B, csf_frac=0.1, zero_frac=0.3):
    L = B + 2
    pw = rng.random((L, L, L))
    pg = 1 - pw
    csf = rng.random((L, L, L)) < csf_frac
    pw[csf] = 0
    pg[csf] = 0
    phi = rng.random((L, L, L)) * 0.6
    phi[rng.random((L, L, L)) < zero_frac] = 0
    return phi, pw, pg


def bits(a):
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def same(a, b):
    return bool(np.all((bits(a) == bits(b)) | (np.isnan(a) & np.isnan(b))))


need_ref = pytest.mark.skipif(not R.available("libgqref_fk_f64.so"), reason="oracle/_ref not built")


@need_ref
@pytest.mark.parametrize("real", ["f64", "f32"])
@pytest.mark.parametrize("B", [4, 16])
def test_fk_block_bit_exact_vs_reference(real, B):
    rng = np.random.default_rng(11 + B)
    dt = np.float64 if real == "f64" else np.float32
    for trial in range(3):
        phi, pw, pg = (a.astype(dt) for a in random_lab(rng, B))
        mine = O.fk_block(phi, pw, pg, H, DW, DG, RHO, mode=real)
        ref = R.fk_block(phi, pw, pg, H, DW, DG, RHO, real=real)
        assert same(mine, ref)
        assert np.isnan(ref).any()  # the CSF quirk is exercised
        p0 = np.ascontiguousarray(phi[1:-1, 1:-1, 1:-1])
        assert same(O.fk_update(p0, mine, 1.7, mode=real), R.update(p0, ref, 1.7, real=real))


@need_ref
def test_fk_block_special_values_vs_reference():
    """Dw = 0 (df = 0 everywhere -> Neumann branch), negative tissue, all-CSF."""
    rng = np.random.default_rng(5)
    B = 4
    phi, pw, pg = random_lab(rng, B)
    pw2 = pw.copy()
    pw2[2, 2, 2] = -0.3  # negative interpolated tissue
    for args in [(0.0, 0.0, RHO), (DW, DG, 0.0), (DW, DG, RHO)]:
        for real in ("f64", "f32"):
            dt = np.float64 if real == "f64" else np.float32
            a = [x.astype(dt) for x in (phi, pw2, pg)]
            assert same(O.fk_block(*a, H, *args, mode=real), R.fk_block(*a, H, *args, real=real))


def test_fk_golden_vectors():
    path = os.path.join(GOLD, "fk_operator.npz")
    g = np.load(path)
    for real in ("f64", "f32"):
        dt = np.float64 if real == "f64" else np.float32
        phi, pw, pg = (g["lab_" + k].astype(dt) for k in ("phi", "pw", "pg"))
        out = O.fk_block(phi, pw, pg, float(g["h"]), float(g["Dw"]), float(g["Dg"]), float(g["rho"]), mode=real)
        assert same(out, g["dphidt_" + real])
        up = O.fk_update(np.ascontiguousarray(phi[1:-1, 1:-1, 1:-1]), out, float(g["dt"]), mode=real)
        assert same(up, g["phi1_" + real])


def test_pure_reaction_is_euler_recurrence():
    """Dw = Dg = 0: every cell follows phi <- clamp(phi + dt*rho*phi*(1-phi))."""
    rng = np.random.default_rng(1)
    N, B = 32, 16
    phi = rng.random((N, N, N)) * 0.5
    pw = rng.random((N, N, N))
    pg = 1 - pw
    dt, steps = 0.7, 5
    out = O.fk_uniform_steps(phi, pw, pg, B, 1 / N, 0.0, 0.0, RHO, dt, steps)
    exp = phi.copy()
    for _ in range(steps):
        exp = np.clip(exp + dt * (RHO * exp * (1. - exp)), 0, 1)
    assert np.array_equal(out, exp)


def test_fourier_mode_decay_uniform_tissue():
    """Uniform white matter (K = Dw): phi = a + e*cos(2 pi k x) decays by
    exactly (1 - dt*Dw*ih2*4 sin^2(pi k h)) per step (rho = 0, no clamp)."""
    N, B, k = 32, 16, 3
    h = 1.0 / N
    x = np.arange(N) * h
    a, e = 0.5, 0.1
    phi = np.broadcast_to(a + e * np.cos(2 * np.pi * k * x)[None, None, :], (N, N, N)).copy()
    pw = np.ones((N, N, N))
    pg = np.zeros((N, N, N))
    Dw = 1e-3
    dt = 0.4 * h * h / (6 * Dw)
    steps = 10
    out = O.fk_uniform_steps(phi, pw, pg, B, h, Dw, 0.1 * Dw, 0.0, dt, steps)
    g = 1 - dt * Dw / (h * h) * 4 * np.sin(np.pi * k * h) ** 2
    exp = a + e * g ** steps * np.cos(2 * np.pi * k * x)
    assert np.max(np.abs(out - exp[None, None, :])) < 1e-12


def test_diffusion_conserves_mass_in_closed_region():
    """rho = 0, tissue confined by CSF-free interior: the diffusive increments
    telescope, so sum(phi) is conserved while phi stays away from the clamp."""
    rng = np.random.default_rng(3)
    N, B = 32, 16
    pw = 0.2 + 0.8 * rng.random((N, N, N))
    pg = 1 - pw
    z, y, x = np.meshgrid(*(np.arange(N),) * 3, indexing="ij")
    phi = 0.4 * np.exp(-((x - 16.) ** 2 + (y - 16.) ** 2 + (z - 16.) ** 2) / 20.)
    h = 1 / N
    Dw = 1e-3
    dt = 0.5 * h * h / (6 * Dw)
    out = O.fk_uniform_steps(phi, pw, pg, B, h, Dw, 0.1 * Dw, 0.0, dt, 20)
    assert out.min() > 0 and out.max() < 1
    assert abs(out.sum() - phi.sum()) < 1e-9 * phi.sum()
    assert np.abs(out - phi).max() > 1e-4  # it did diffuse


def test_csf_neighbour_forces_zero():
    """A tissue cell next to a CSF cell gets D/b = 0/0 -> NaN -> phi = 0 after
    the update; a cell with no tissue and no tumour stays 0."""
    B = 4
    L = B + 2
    phi = np.full((L, L, L), 0.3)
    pw = np.full((L, L, L), 0.7)
    pg = np.full((L, L, L), 0.3)
    pw[3, 3, 3] = pg[3, 3, 3] = 0.0  # CSF at lab (3,3,3) = block (2,2,2)
    d = O.fk_block(phi, pw, pg, H, DW, DG, RHO)
    p1 = O.fk_update(phi[1:-1, 1:-1, 1:-1], d, 1.0)
    # face neighbours of the CSF cell and the CSF cell itself (tissue = phi > 0)
    for c in [(2, 2, 2), (1, 2, 2), (3, 2, 2), (2, 1, 2), (2, 3, 2), (2, 2, 1), (2, 2, 3)]:
        assert p1[c] == 0.0
    assert np.count_nonzero(p1 == 0.0) == 7


def test_harmonic_mean_symmetry():
    """Swapping the roles of two face neighbours permutes df symmetrically: a
    mirrored lab gives the mirrored dphidt."""
    rng = np.random.default_rng(9)
    phi, pw, pg = random_lab(rng, 4, csf_frac=0.0)
    d = O.fk_block(phi, pw, pg, H, DW, DG, RHO)
    dm = O.fk_block(phi[:, :, ::-1].copy(), pw[:, :, ::-1].copy(), pg[:, :, ::-1].copy(), H, DW, DG, RHO)
    np.testing.assert_allclose(dm[:, :, ::-1], d, rtol=1e-13, atol=1e-18)


@need_ref
@pytest.mark.parametrize("real", ["f32", "f64"])
def test_uniform_steps_bit_exact_vs_reference_driver(real):
    """Several steps of the oracle's uniform driver against the reference's
    functors driven block by block with CCell_extended-sized cells (the CPU
    baseline of the benchmark), on anatomy with CSF and a tumour."""
    rng = np.random.default_rng(5)
    N = 32
    phi = np.zeros((N, N, N))
    phi[12:20, 10:18, 14:22] = rng.random((8, 8, 8)) * 0.8
    pw = rng.random((N, N, N))
    pg = 1 - pw
    csf = rng.random((N, N, N)) < 0.05
    pw[csf] = 0
    pg[csf] = 0
    h = 1.0 / N
    dt = 0.9 * h * h / (6 * DW)
    mode = "f64" if real == "f64" else "f32ref"
    dtp = np.float64 if real == "f64" else np.float32
    a = [x.astype(dtp) for x in (phi, pw, pg)]
    want = R.uniform_steps(a[0], a[1], a[2], h, DW, DG, RHO, dt, 6, threads=4, real=real)
    got = O.fk_uniform_steps(a[0], a[1], a[2], 16, h, DW, DG, RHO, dt, 6, mode=mode)
    assert (want > 0).sum() > 1000
    assert same(np.ascontiguousarray(got, dtp), want)
