"""GPU surrogate (gq_sur_*): bf16 MFMA forward/backward against the CPU
oracle (oracle/surrogate_oracle.py, pinned by finite differences in
test_oracle_surrogate.py) on identical parameters and inputs.

Bars: the oracle runs with the same bf16 rounding points in float64, so the
remaining differences are FP32 accumulation order and the occasional
one-ulp bf16 flip it causes: prediction relative L2 <= 2e-3; every layer's
gradient relative L2 <= 2e-3 and cosine >= 0.99999; the loss within 1e-5
relative.  Training (Adam) on a fixed batch lowers the loss; results are
bit-reproducible.
"""
import numpy as np
import pytest

from gliomauq_amd import surrogate as S
from oracle import surrogate_oracle as SO

pytestmark = pytest.mark.gpu


def batch(seed=0, B=1):
    rng = np.random.default_rng(seed)
    n = 64
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    pw = (0.5 + 0.4 * np.sin(x / 7.0) * np.cos(y / 5.0)).astype(np.float32)
    pg = (1 - pw).astype(np.float32)
    params = rng.uniform(-1, 1, (B, 4)).astype(np.float32)
    u0 = np.zeros((B, n, n, n), np.float32)
    tgt = np.zeros((B, n, n, n), np.float32)
    for b in range(B):
        c = rng.uniform(20, 44, 3)
        r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
        u0[b] = np.exp(-r2 / 8.0)
        tgt[b] = 1.0 / (1.0 + np.exp((np.sqrt(r2) - 13.0 - 3 * params[b, 0]) / 1.5))
    return params, u0, pw, pg, tgt


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_surrogate_matches_oracle():
    params, u0, pw, pg, tgt = batch()
    s = S.Surrogate(batch=2, seed=3)
    s.set_anatomy(pw, pg)
    theta = s.get_params().astype(np.float64)
    pred = s.predict(params, u0)
    loss, g = s.loss_grad(params, u0, tgt)
    l_ref, p_ref, g_ref = SO.forward_backward(theta, params.astype(np.float64), u0.astype(np.float64),
                                             pw.astype(np.float64), pg.astype(np.float64), tgt.astype(np.float64))
    e_p = rel(pred, p_ref)
    print("prediction rel L2 %.3e, loss gpu %.6e oracle %.6e" % (e_p, loss, l_ref))
    assert e_p <= 2e-3
    assert abs(loss - l_ref) <= 1e-5 * abs(l_ref)
    o = 0
    for i, (cin, cout) in enumerate(SO.SPEC + [(16, 1)]):
        n = cout * 27 * cin + cout if i < 7 else 17
        a, b = g[o:o + n].astype(np.float64), g_ref[o:o + n]
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-300))
        print("layer %d grad rel L2 %.3e cos %.6f" % (i, rel(a, b), cos))
        assert rel(a, b) <= 2e-3 and cos >= 0.99999
        o += n
    s.close()


def test_surrogate_adam_steps_match_oracle():
    """Three Adam steps on the device against Adam on the oracle's gradients
    from the same starting parameters (the first Adam step moves every
    parameter by ~lr * sign(g), so a wrong or missing update shows at once)."""
    params, u0, pw, pg, tgt = batch(seed=1, B=1)
    s = S.Surrogate(batch=1, seed=4)
    s.set_anatomy(pw, pg)
    th = s.get_params().astype(np.float64)
    m = np.zeros_like(th)
    v = np.zeros_like(th)
    lr = 1e-3
    for it in range(3):
        s.train_step(params, u0, tgt, lr)
        _, _, g = SO.forward_backward(th, params.astype(np.float64), u0.astype(np.float64), pw.astype(np.float64),
                                      pg.astype(np.float64), tgt.astype(np.float64))
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        th = th - lr * (m / (1 - 0.9 ** (it + 1))) / (np.sqrt(v / (1 - 0.999 ** (it + 1))) + 1e-8)
    got = s.get_params().astype(np.float64)
    moved = np.abs(got - (th + 0))
    step = 3 * lr
    frac_bad = float(np.mean(moved > 0.5 * step))
    print("params off by more than half the 3-step budget: %.4f%%" % (100 * frac_bad))
    assert frac_bad < 1e-3  # only sign flips of near-zero gradients may disagree
    s.close()


def test_surrogate_training_and_reproducibility():
    params, u0, pw, pg, tgt = batch(seed=1, B=2)
    s = S.Surrogate(batch=2, seed=4)
    s.set_anatomy(pw, pg)
    p1 = s.predict(params, u0)
    p2 = s.predict(params, u0)
    assert np.array_equal(p1, p2)
    losses = [s.train_step(params, u0, tgt, 3e-4) for _ in range(120)]
    mse = float(np.mean((s.predict(params, u0) - tgt) ** 2))
    # well below what a constant prediction can reach (the tumours fill ~5% of the volume)
    const = min(np.mean((c - tgt) ** 2) for c in np.linspace(0, 0.3, 31))
    print("loss %.4e -> %.4e, prediction MSE %.4e (best constant %.4e)" % (losses[0], losses[-1], mse, const))
    assert losses[-1] < 0.5 * losses[0]
    assert mse < 0.5 * const
    t = s.timing()
    assert t["forward_ms"] > 0 and t["backward_ms"] > 0
    s.close()


def test_surrogate_rejects_bad_batch():
    s = S.Surrogate(batch=1, seed=0)
    params, u0, pw, pg, tgt = batch(B=2)
    s.set_anatomy(pw, pg)
    with pytest.raises(S.GQError) as e:
        s.train_step(params, u0, tgt, 1e-3)
    assert e.value.status == 1
    s.close()


def test_surrogate_likelihood_matches_oracle():
    from oracle import oracle as O
    params, u0, pw, pg, tgt = batch(seed=2, B=2)
    s = S.Surrogate(batch=2, seed=5)
    s.set_anatomy(pw, pg)
    rng = np.random.default_rng(0)
    t1 = (rng.random((64, 64, 64)) < 0.3).astype(np.uint8)
    t2 = (rng.random((64, 64, 64)) < 0.5).astype(np.uint8)
    pet = np.where(rng.random((64, 64, 64)) < 0.2, rng.random((64, 64, 64)), np.nan).astype(np.float32)
    s.set_observations(t1, t2, pet)
    ll = s.loglik_flip(params, u0)
    pred = s.predict(params, u0)
    for b in range(2):
        ref, parts = O.lik_flip(pred[b].astype(np.float64), t1, t2, pet)
        assert abs(ll[b, 0] - ref) <= 1e-10 * abs(ref)
        assert np.allclose(ll[b, 1:], parts, rtol=1e-10)
    s.close()
