"""CPU oracle of the 3D convolutional surrogate -- TEST INFRASTRUCTURE ONLY.

The reference has no learned model; this restates the network that
gliomauq_amd/csrc/surrogate_api.cpp trains (same layers, parameter layout,
loss) in float64 numpy, forward and backward, so the device's bf16-MFMA
kernels can be checked against it.  With bf16=True the activations, weights
and activation gradients are rounded to bf16 (round to nearest even) at the
points where the device stores them in bf16; with bf16=False everything is
float64 (used to pin the backward pass by finite differences).

Layout: channels-last arrays x[b, z, y, x, c]; a 3^3 weight tensor is
[cout][27*cin] with k = tap*cin + ci, tap = (dz+1)*9 + (dy+1)*3 + (dx+1);
zero padding.
"""
import numpy as np

SPEC = [(8, 16), (16, 32), (32, 64), (64, 128), (192, 64), (96, 32), (48, 16)]  # E1 E2 E3 E4 D3 D2 D1
TH, KAPPA, LAM = 0.25, 20.0, 0.1
LEAK = 0.05  # leaky ReLU slope


def bf16(x):
    """Round to bf16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    r = ((b >> 16) & 1) + 0x7FFF
    b = ((b + r) & 0xFFFF0000).astype(np.uint32)
    return b.view(np.float32).astype(np.float64)


def unpack(theta):
    """Flat parameter vector -> list of (W [cout][27*cin], b) and head (w[16], b)."""
    out = []
    o = 0
    for cin, cout in SPEC:
        n = cout * 27 * cin
        w = theta[o:o + n].reshape(cout, 27 * cin)
        o += n
        b = theta[o:o + cout]
        o += cout
        out.append((w, b))
    return out, (theta[o:o + 16], theta[o + 16])


def n_params():
    return sum(c * 27 * ci + c for ci, c in SPEC) + 17


def _taps():
    for tap in range(27):
        yield tap, tap % 3 - 1, (tap // 3) % 3 - 1, tap // 9 - 1


def conv3(x, w, b=None):
    """x [B,D,D,D,cin], w [cout][27*cin] -> [B,D,D,D,cout] (zero padding)."""
    B, D = x.shape[0], x.shape[1]
    cin = x.shape[4]
    cout = w.shape[0]
    xp = np.zeros((B, D + 2, D + 2, D + 2, cin))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    y = np.zeros((B, D, D, D, cout))
    for tap, dx, dy, dz in _taps():
        xs = xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + D, 1 + dx:1 + dx + D]
        y += xs @ w[:, tap * cin:(tap + 1) * cin].T
    if b is not None:
        y += b
    return y


def conv3_wgrad(x, dz):
    """dW[co][tap*cin+ci] = sum dz[v][co] x[v+tap][ci]"""
    B, D = x.shape[0], x.shape[1]
    cin = x.shape[4]
    cout = dz.shape[4]
    xp = np.zeros((B, D + 2, D + 2, D + 2, cin))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    dw = np.zeros((cout, 27 * cin))
    dzf = dz.reshape(-1, cout)
    for tap, dx, dy, dz_ in _taps():
        xs = xp[:, 1 + dz_:1 + dz_ + D, 1 + dy:1 + dy + D, 1 + dx:1 + dx + D].reshape(-1, cin)
        dw[:, tap * cin:(tap + 1) * cin] = dzf.T @ xs
    return dw


def conv3_bwd_data(dz, w, cin):
    """dX[v][ci] = sum_tap sum_co dz[v - tap][co] W[co][tap*cin + ci]"""
    B, D = dz.shape[0], dz.shape[1]
    cout = dz.shape[4]
    dzp = np.zeros((B, D + 2, D + 2, D + 2, cout))
    dzp[:, 1:-1, 1:-1, 1:-1] = dz
    dx = np.zeros((B, D, D, D, cin))
    for tap, ox, oy, oz in _taps():
        ds = dzp[:, 1 - oz:1 - oz + D, 1 - oy:1 - oy + D, 1 - ox:1 - ox + D]
        dx += ds @ w[:, tap * cin:(tap + 1) * cin]
    return dx


def pool(x):
    B, D, C = x.shape[0], x.shape[1], x.shape[4]
    return x.reshape(B, D // 2, 2, D // 2, 2, D // 2, 2, C).mean(axis=(2, 4, 6))


def pool_bwd(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_bwd(dy):
    B, D, C = dy.shape[0], dy.shape[1], dy.shape[4]
    return dy.reshape(B, D // 2, 2, D // 2, 2, D // 2, 2, C).sum(axis=(2, 4, 6))


def assemble(params, u0, pw, pg):
    """params [B,4], u0 [B,D,D,D] ([z,y,x]), pw/pg [D,D,D] -> x [B,D,D,D,8]"""
    B, D = u0.shape[0], u0.shape[1]
    x = np.zeros((B, D, D, D, 8))
    x[..., 0:4] = params[:, None, None, None, :]
    x[..., 4] = u0
    x[..., 5] = pw
    x[..., 6] = pg
    x[..., 7] = 1.0
    return x


def forward_backward(theta, params, u0, pw, pg, target=None, bf=True):
    """Loss, prediction and (with a target) the gradient w.r.t. theta."""
    R = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    convs, (hw, hb) = unpack(np.asarray(theta, np.float64))
    W = [R(w) for w, _ in convs]
    Bs = [b for _, b in convs]
    relu = lambda a: np.where(a > 0, a, LEAK * a)  # noqa: E731
    x0 = R(assemble(params, u0, pw, pg))
    e1 = R(relu(conv3(x0, W[0], Bs[0])))
    p1 = R(pool(e1))
    e2 = R(relu(conv3(p1, W[1], Bs[1])))
    p2 = R(pool(e2))
    e3 = R(relu(conv3(p2, W[2], Bs[2])))
    p3 = R(pool(e3))
    e4 = R(relu(conv3(p3, W[3], Bs[3])))
    c3 = np.concatenate([up(e4), e3], 4)
    d3 = R(relu(conv3(c3, W[4], Bs[4])))
    c2 = np.concatenate([up(d3), e2], 4)
    d2 = R(relu(conv3(c2, W[5], Bs[5])))
    c1 = np.concatenate([up(d2), e1], 4)
    d1 = R(relu(conv3(c1, W[6], Bs[6])))
    zl = d1 @ hw + hb
    p = 1.0 / (1.0 + np.exp(-zl))
    if target is None:
        return None, p, None
    n = p.size
    m = (target > TH).astype(np.float64)
    s = 1.0 / (1.0 + np.exp(-KAPPA * (p - TH)))
    sc = np.clip(s, 1e-7, 1 - 1e-7)
    loss = np.sum((p - target) ** 2 + LAM * -(m * np.log(sc) + (1 - m) * np.log(1 - sc))) / n
    dldp = (2 * (p - target) + LAM * (s - m) * KAPPA) / n
    dzl = dldp * p * (1 - p)
    g_hw = np.einsum("bzyxc,bzyx->c", d1, dzl)
    g_hb = dzl.sum()
    dd1 = dzl[..., None] * hw
    grads = [None] * 7

    def layer(i, dy, y, x):
        dz = R(dy * np.where(y > 0, 1.0, LEAK))
        grads[i] = (conv3_wgrad(x, dz), dz.sum(axis=(0, 1, 2, 3)))
        return conv3_bwd_data(dz, W[i], x.shape[4])

    dc1 = layer(6, dd1, d1, c1)
    dd2 = up_bwd(dc1[..., :32])
    dc2 = layer(5, dd2, d2, c2)
    dd3 = up_bwd(dc2[..., :64])
    dc3 = layer(4, dd3, d3, c3)
    de4 = up_bwd(dc3[..., :128])
    dp3 = layer(3, de4, e4, p3)
    de3 = dc3[..., 128:] + pool_bwd(dp3)
    dp2 = layer(2, de3, e3, p2)
    de2 = dc2[..., 64:] + pool_bwd(dp2)
    dp1 = layer(1, de2, e2, p1)
    de1 = dc1[..., 32:] + pool_bwd(dp1)
    layer(0, de1, e1, x0)
    flat = []
    for gw, gb in grads:
        flat += [gw.ravel(), gb.ravel()]
    flat += [g_hw, np.array([g_hb])]
    return loss, p, np.concatenate(flat)
