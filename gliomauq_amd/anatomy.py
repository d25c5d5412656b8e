"""Seeded synthetic anatomy, initial condition and observations.

Stands in for the patient volumes the reference reads in
Glioma_HG_UQ::_ic_PatientCase (Glioma/Glioma_HG_UQ.cpp:67-225), which are not
part of the reference tree.  Everything is defined on the uniform vertex
grid of n^3 points over a cube of edge L cm (point i at x = i*L/n), arrays
indexed [z, y, x]:

  * brain: ellipsoid with semi-axes (7.0, 6.0, 5.5) cm centred in the cube;
  * two CSF ventricles, ellipsoids of semi-axes (0.45, 0.75, 0.5) cm
    mirrored about the mid-sagittal plane (x = L/2 +- 0.8 cm), where
    Pwm = Pgm = 0;
  * elsewhere in the brain Pwm = sigmoid(8 g), Pgm = 1 - Pwm, g a unit-variance
    Gaussian random field of correlation length 1 cm (white noise smoothed by
    a Gaussian kernel, FFT, periodic);
  * outside the brain Pwm = Pgm = 0;
  * initial condition: Gaussian bump, amplitude 0.1, sigma 2 voxels, centred
    on a seeded voxel with Pwm > 0.5 at least 1.5 cm inside the brain.

Observations (flip-rate model): from a ground-truth field u, T1Gd = [u > 0.7]
and FLAIR = [u > 0.25] with every voxel flipped with probability 0.05, and
PET = u + N(0, 0.05^2) on voxels with u > 0.25 (NaN elsewhere).
"""
import numpy as np

BRAIN_AXES = (7.0, 6.0, 5.5)
VENT_AXES = (0.45, 0.75, 0.5)
VENT_DX = 0.8


def coords(n, L):
    x = np.arange(n) * (L / n)
    return np.meshgrid(x, x, x, indexing="ij")  # z, y, x


def anatomy(n, L=17.0, seed=0, corr_cm=1.0):
    """Returns (pwm, pgm) float32 [z,y,x] and the brain mask."""
    rng = np.random.default_rng(seed)
    z, y, x = coords(n, L)
    c = L / 2
    brain = ((x - c) / BRAIN_AXES[0]) ** 2 + ((y - c) / BRAIN_AXES[1]) ** 2 + ((z - c) / BRAIN_AXES[2]) ** 2 <= 1.0
    vent = np.zeros_like(brain)
    for s in (-1, 1):
        vent |= (((x - c - s * VENT_DX) / VENT_AXES[0]) ** 2 + ((y - c) / VENT_AXES[1]) ** 2 +
                 ((z - c - 0.3) / VENT_AXES[2]) ** 2) <= 1.0
    # Gaussian random field, correlation length corr_cm (Gaussian kernel, FFT)
    noise = rng.standard_normal((n, n, n))
    k = np.fft.fftfreq(n, d=L / n) * 2 * np.pi
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    filt = np.exp(-0.5 * (kx ** 2 + ky ** 2 + kz ** 2) * corr_cm ** 2)
    g = np.real(np.fft.ifftn(np.fft.fftn(noise) * filt))
    g = (g - g.mean()) / g.std()
    pwm = 1.0 / (1.0 + np.exp(-8.0 * g))
    tissue = brain & ~vent
    pwm = np.where(tissue, pwm, 0.0)
    pgm = np.where(tissue, 1.0 - pwm, 0.0)
    return pwm.astype(np.float32), pgm.astype(np.float32), tissue


def pick_seed_voxel(pwm, n, L=17.0, seed=0, margin_cm=1.5):
    rng = np.random.default_rng(seed + 1000)
    z, y, x = coords(n, L)
    c = L / 2
    inner = (((x - c) / (BRAIN_AXES[0] - margin_cm)) ** 2 + ((y - c) / (BRAIN_AXES[1] - margin_cm)) ** 2 +
             ((z - c) / (BRAIN_AXES[2] - margin_cm)) ** 2) <= 1.0
    cand = np.argwhere(inner & (pwm > 0.5))
    return tuple(int(v) for v in cand[rng.integers(len(cand))])  # (iz, iy, ix)


def initial_condition(n, center_zyx, amplitude=0.1, sigma_vox=2.0):
    """Gaussian bump (periodic distance) on the n^3 grid, float32 [z,y,x]."""
    i = np.arange(n)
    d = [np.minimum(np.abs(i - c), n - np.abs(i - c)).astype(np.float64) for c in center_zyx]
    dz, dy, dx = np.meshgrid(d[0], d[1], d[2], indexing="ij")
    u = amplitude * np.exp(-(dx ** 2 + dy ** 2 + dz ** 2) / (2 * sigma_vox ** 2))
    return u.astype(np.float32)


def observations(u, seed=0, th1=0.7, th2=0.25, flip=0.05, sigma=0.05):
    """(t1 u8, flair u8, pet float32 with NaN) from a ground-truth field u."""
    rng = np.random.default_rng(seed + 2000)
    m1 = (u > th1)
    m2 = (u > th2)
    f1 = rng.random(u.shape) < flip
    f2 = rng.random(u.shape) < flip
    t1 = (m1 ^ f1).astype(np.uint8)
    t2 = (m2 ^ f2).astype(np.uint8)
    pet = np.where(m2, u + sigma * rng.standard_normal(u.shape), np.nan).astype(np.float32)
    return t1, t2, pet


def config_inputs(n, L=17.0, seed=0):
    """(pwm, pgm, u0, seed voxel) for the synthetic-anatomy workloads."""
    pwm, pgm, _ = anatomy(n, L, seed)
    c = pick_seed_voxel(pwm, n, L, seed)
    u0 = initial_condition(n, c)
    return pwm, pgm, u0, c


def ensemble_prior(S, n, L=17.0, seed=0, pwm=None, cfl=0.9):
    """The batched-ensemble workload's samples: S parameter sets from a seeded
    prior -- log-uniform Dw in [2e-4, 1.5e-2] cm^2/day, log-uniform rho in
    [2e-3, 0.2] /day, initial point uniform over the voxels with Pwm > 0.5, T
    uniform in [30, 300] days -- as (params (S, 4) = Dw, Dg, rho, dt in
    unit-cube units with dt at `cfl` of the finest level of n^3, step counts of
    `while (t <= T)`, Gaussian ICs (S, 5) = x0, y0, z0, 0.1, 2 voxels)."""
    if pwm is None:
        pwm = anatomy(n, L, seed)[0]
    rng = np.random.default_rng(seed + 77)
    wm = np.argwhere(pwm > 0.5)
    Dw_cm = np.exp(rng.uniform(np.log(2e-4), np.log(1.5e-2), S))
    rho = np.exp(rng.uniform(np.log(2e-3), np.log(0.2), S))
    T = rng.uniform(30.0, 300.0, S)
    centers = wm[rng.integers(len(wm), size=S)]  # (z, y, x)
    h = 1.0 / n
    prm = np.zeros((S, 4))
    steps = np.zeros(S, np.int32)
    for s in range(S):
        Dw = Dw_cm[s] / L ** 2
        dt = cfl * h * h / (6 * Dw)
        prm[s] = (Dw, 0.1 * Dw, rho[s], dt)
        t, k = 0.0, 0
        while t <= T[s]:
            k += 1
            t += dt
        steps[s] = k
    ic = np.array([[c[2], c[1], c[0], 0.1, 2.0] for c in centers], np.float64)
    return prm, steps, ic
