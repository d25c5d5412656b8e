"""The surrogate oracle (oracle/surrogate_oracle.py) pinned by finite
differences of its own float64 forward pass (bf16 rounding off), on a 16^3
volume (the layers are size-generic; the device network runs at 64^3).

Random directions in parameter space and single coordinates of every layer
are checked, so a dropped skip connection, a wrong pooling factor, a
transposed weight block or a missing ReLU mask fails one of them.  A second
check ties the bf16 rounding helper to round-to-nearest-even.
"""
import numpy as np

from oracle import surrogate_oracle as SO

D = 16


def inputs(seed=0, B=2):
    rng = np.random.default_rng(seed)
    theta = np.empty(SO.n_params())
    o = 0
    for cin, cout in SO.SPEC:
        n = cout * 27 * cin
        theta[o:o + n] = rng.standard_normal(n) * np.sqrt(2.0 / (27 * cin))
        o += n
        theta[o:o + cout] = 0.05 * rng.standard_normal(cout)
        o += cout
    theta[o:o + 16] = 0.3 * rng.standard_normal(16)
    theta[o + 16] = 0.1
    params = rng.uniform(-1, 1, (B, 4))
    u0 = rng.random((B, D, D, D)) * 0.3
    pw = rng.random((D, D, D))
    pg = 1 - pw
    target = np.clip(rng.random((B, D, D, D)) * 0.8 + 0.1 * np.sin(np.arange(D))[None, None, None, :], 0, 1)
    return theta, params, u0, pw, pg, target


def test_layer_adjoint_identities():
    """<conv3(x, W), dz> = <W, wgrad(x, dz)> = <x, bwd_data(dz, W)>, and the
    same pairing for pooling and upsampling (exact bilinear identities)."""
    rng = np.random.default_rng(2)
    for cin, cout in ((8, 16), (48, 16), (32, 64)):
        x = rng.standard_normal((2, 8, 8, 8, cin))
        w = rng.standard_normal((cout, 27 * cin))
        dz = rng.standard_normal((2, 8, 8, 8, cout))
        y = SO.conv3(x, w)
        lhs = np.sum(y * dz)
        assert abs(np.sum(w * SO.conv3_wgrad(x, dz)) - lhs) <= 1e-10 * abs(lhs)
        assert abs(np.sum(x * SO.conv3_bwd_data(dz, w, cin)) - lhs) <= 1e-10 * abs(lhs)
    x = rng.standard_normal((2, 8, 8, 8, 3))
    g = rng.standard_normal((2, 4, 4, 4, 3))
    assert abs(np.sum(SO.pool(x) * g) - np.sum(x * SO.pool_bwd(g))) <= 1e-12 * abs(np.sum(SO.pool(x) * g))
    G = rng.standard_normal((2, 16, 16, 16, 3))
    assert abs(np.sum(SO.up(x) * G) - np.sum(x * SO.up_bwd(G))) <= 1e-12 * abs(np.sum(SO.up(x) * G))
    # a conv tap: a unit input at the centre reaches the output at -offset
    x = np.zeros((1, 5, 5, 5, 8))
    x[0, 2, 2, 2, 3] = 1.0
    w = np.zeros((16, 27 * 8))
    tap = (1 + 1) * 9 + (0 + 1) * 3 + (-1 + 1)  # dz=+1, dy=0, dx=-1
    w[5, tap * 8 + 3] = 2.0
    y = SO.conv3(x, w)
    assert y[0, 1, 2, 3, 5] == 2.0 and np.count_nonzero(y) == 1


def test_oracle_gradient_matches_finite_differences():
    theta, params, u0, pw, pg, target = inputs()
    loss, p, g = SO.forward_backward(theta, params, u0, pw, pg, target, bf=False)
    assert np.isfinite(loss) and g.shape == theta.shape

    def L(t):
        return SO.forward_backward(t, params, u0, pw, pg, target, bf=False)[0]

    rng = np.random.default_rng(1)
    o = 0
    for cin, cout in SO.SPEC:
        n = cout * 27 * cin
        # every bias of the layer (a direction through the whole backward chain)
        v = np.zeros_like(theta)
        v[o + n:o + n + cout] = rng.standard_normal(cout)
        fd = (L(theta + 1e-6 * v) - L(theta - 1e-6 * v)) / 2e-6
        assert abs(fd - g @ v) <= 1e-6 * abs(fd) + 1e-12, (o, fd, g @ v)
        # single weights (few ReLU switches inside the difference quotient)
        for i in rng.integers(n, size=2):
            e = np.zeros_like(theta)
            e[o + i] = 1e-6
            fd = (L(theta + e) - L(theta - e)) / 2e-6
            assert abs(fd - g[o + i]) <= 1e-4 * max(abs(fd), 1e-6), (o + i, fd, g[o + i])
        o += n + cout
    v = np.zeros_like(theta)
    v[o:o + 17] = rng.standard_normal(17)
    fd = (L(theta + 1e-6 * v) - L(theta - 1e-6 * v)) / 2e-6
    assert abs(fd - g @ v) <= 1e-6 * abs(fd)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 3.0e-39])
    r = SO.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0 + 2 ** -7 or r[1] == 1.0  # 2^-8 is half an ulp (ulp 2^-7): ties to even -> 1.0
    assert r[1] == 1.0
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == 1.0
    assert r[4] == -2.5
