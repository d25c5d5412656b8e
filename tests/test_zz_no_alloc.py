"""gq_loglik_flip_volume_noalloc -- the flip-rate log-likelihood of a model
volume in device memory with no allocation (its partial sums in the
library's fixed device array) -- against gq_loglik_flip on the same field and
observations: bit-identical, on two calls in a row, the second on a smaller
model grid with a coarser observation grid.  The test's own device buffers
come from the HIP runtime libgq.so runs on; the calls go through the C-ABI."""
import ctypes as C

import numpy as np
import pytest

import gliomauq_amd as gq
from gliomauq_amd import _lib

pytestmark = pytest.mark.gpu

PARAMS = np.array([0.7, 0.25, 0.05, 0.05], np.float64)


def grid_field(ctx, bpd, seed):
    rng = np.random.default_rng(seed)
    n = bpd * 16
    g = gq.Grid(ctx, bpd, int(round(np.log2(bpd))), 4)
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    phi = np.exp(-((x - n / 2) ** 2 + (y - n / 3) ** 2 + (z - n / 2) ** 2) / (0.05 * n * n))
    phi = (phi + 0.02 * rng.random(phi.shape)).astype(np.float32)
    g.sample_volume("phi", phi)
    return g


def observations(n, seed):
    rng = np.random.default_rng(seed + 100)
    t1 = (rng.random((n, n, n)) < 0.2).astype(np.uint8)
    t2 = (rng.random((n, n, n)) < 0.4).astype(np.uint8)
    pet = np.where(rng.random((n, n, n)) < 0.3, rng.random((n, n, n)), np.nan).astype(np.float32)
    return t1, t2, pet


def hip():
    """The HIP runtime libgq.so runs on (the same loaded library)."""
    _lib.load()
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipDeviceSynchronize.argtypes = []
    return H


def to_device(H, a):
    a = np.ascontiguousarray(a)
    p = C.c_void_p()
    assert H.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
    assert H.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
    return p


def noalloc(model, t1, t2, pet):
    H = hip()
    bufs = [to_device(H, v) for v in (model, t1, t2, pet, np.full(4, np.nan))]
    try:
        L = _lib.load()
        st = L.gq_loglik_flip_volume_noalloc(bufs[0], model.shape[0], bufs[1], bufs[2], bufs[3], t1.shape[0],
                                             _lib.ptr(PARAMS, C.c_double), bufs[4], None)
        assert st == 0, st
        assert H.hipDeviceSynchronize() == 0
        out = np.zeros(4)
        assert H.hipMemcpy(out.ctypes.data, bufs[4], 32, 2) == 0  # device to host
        return out
    finally:
        for b in bufs:
            H.hipFree(b)


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_matches_loglik_flip(order):
    ctx = gq.Context(0)
    for bpd, obs_n, seed in ((8, 128, 1), (2, 16, 2)):
        g = grid_field(ctx, bpd, seed)
        t1, t2, pet = observations(obs_n, seed)
        want = gq.FlipLikelihood(g, t1, t2, pet).loglik()
        model = g.dump_uniform()
        got = noalloc(model, t1, t2, pet)
        print("gpd %d obs %d: noalloc %r, gq_loglik_flip %r" % (model.shape[0], obs_n, got.tolist(), want.tolist()))
        assert got.tobytes() == np.asarray(want, np.float64).tobytes()
    ctx.close()


def test_noalloc_rejects_bad_sizes():
    L = _lib.load()
    buf = np.zeros(8)
    p = _lib.ptr(PARAMS, C.c_double)
    vp = C.c_void_p(buf.ctypes.data)
    # sizes beyond the maximum, observation grids that do not divide the model grid, bad parameters
    assert L.gq_loglik_flip_volume_noalloc(vp, 2048, vp, vp, vp, 2048, p, vp, None) == 1
    assert L.gq_loglik_flip_volume_noalloc(vp, 64, vp, vp, vp, 48, p, vp, None) == 1
    assert L.gq_loglik_flip_volume_noalloc(None, 64, vp, vp, vp, 64, p, vp, None) == 1
    bad = np.array([0.7, 0.25, 1.5, 0.05])
    assert L.gq_loglik_flip_volume_noalloc(vp, 64, vp, vp, vp, 64, _lib.ptr(bad, C.c_double), vp, None) == 7
