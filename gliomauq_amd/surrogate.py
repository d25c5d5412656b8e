"""Python side of the learned surrogate (gq_sur_* in libgq.so).

A 3D convolutional encoder-decoder on 64^3 that maps four parameters
(broadcast as channels), the initial condition u0 and the anatomy (pw, pg)
to the final tumour field, trained and served by the project's bf16 MFMA
kernels (FP32 accumulation, FP32 master weights, Adam).  It is the first
stage of delayed-acceptance MCMC (sampler.da_rwm); the full solver judges
every proposal it lets through.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import GQError, ptr

N = 64


def _sig(L):
    vp = C.c_void_p
    f32p = C.POINTER(C.c_float)
    f64p = C.POINTER(C.c_double)
    table = {
        "gq_sur_create": [C.POINTER(vp), C.c_int, C.c_int, C.c_uint64],
        "gq_sur_destroy": [vp],
        "gq_sur_set_anatomy": [vp, f32p, f32p],
        "gq_sur_param_count": [vp, C.POINTER(C.c_int64)],
        "gq_sur_get_params": [vp, f32p],
        "gq_sur_set_params": [vp, f32p],
        "gq_sur_predict": [vp, f32p, f32p, C.c_int, f32p],
        "gq_sur_loss_grad": [vp, f32p, f32p, f32p, C.c_int, f64p, f32p],
        "gq_sur_train_step": [vp, f32p, f32p, f32p, C.c_int, C.c_float, f64p],
        "gq_sur_timing": [vp, f64p],
        "gq_sur_obs_flip_upload": [vp, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), f32p],
        "gq_sur_loglik_flip": [vp, f32p, f32p, C.c_int, f64p, f64p],
    }
    for name, args in table.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.gq_sur_last_error.argtypes = [vp]
    L.gq_sur_last_error.restype = C.c_char_p
    return L


class Surrogate:
    def __init__(self, batch=8, seed=0, device=0):
        self.L = _sig(_lib.load())
        h = C.c_void_p()
        st = self.L.gq_sur_create(C.byref(h), device, batch, seed)
        if st != _lib.GQ_OK:
            raise GQError(st, "gq_sur_create")
        self.h = h
        self.batch = batch

    def close(self):
        if self.h:
            self.L.gq_sur_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, name, *args):
        st = getattr(self.L, name)(self.h, *args)
        if st != _lib.GQ_OK:
            raise GQError(st, self.L.gq_sur_last_error(self.h).decode())

    def set_anatomy(self, pw, pg):
        self._keep = [np.ascontiguousarray(a, np.float32) for a in (pw, pg)]
        self._c("gq_sur_set_anatomy", ptr(self._keep[0], C.c_float), ptr(self._keep[1], C.c_float))

    def n_params(self):
        n = C.c_int64()
        self._c("gq_sur_param_count", C.byref(n))
        return n.value

    def get_params(self):
        out = np.zeros(self.n_params(), np.float32)
        self._c("gq_sur_get_params", ptr(out, C.c_float))
        return out

    def set_params(self, theta):
        a = np.ascontiguousarray(theta, np.float32)
        self._c("gq_sur_set_params", ptr(a, C.c_float))

    @staticmethod
    def _arr(params, u0, target=None):
        p = np.ascontiguousarray(params, np.float32).reshape(-1, 4)
        u = np.ascontiguousarray(u0, np.float32).reshape(len(p), N, N, N)
        t = None if target is None else np.ascontiguousarray(target, np.float32).reshape(len(p), N, N, N)
        return p, u, t

    def predict(self, params, u0):
        p, u, _ = self._arr(params, u0)
        out = np.zeros((len(p), N, N, N), np.float32)
        for i in range(0, len(p), self.batch):
            b = min(self.batch, len(p) - i)
            self._c("gq_sur_predict", ptr(p[i:i + b], C.c_float), ptr(u[i:i + b], C.c_float), b,
                    ptr(out[i:i + b], C.c_float))
        return out

    def loss_grad(self, params, u0, target):
        p, u, t = self._arr(params, u0, target)
        g = np.zeros(self.n_params(), np.float32)
        loss = C.c_double()
        self._c("gq_sur_loss_grad", ptr(p, C.c_float), ptr(u, C.c_float), ptr(t, C.c_float), len(p),
                C.byref(loss), ptr(g, C.c_float))
        return loss.value, g

    def train_step(self, params, u0, target, lr):
        p, u, t = self._arr(params, u0, target)
        loss = C.c_double()
        self._c("gq_sur_train_step", ptr(p, C.c_float), ptr(u, C.c_float), ptr(t, C.c_float), len(p),
                C.c_float(lr), C.byref(loss))
        return loss.value

    def set_observations(self, t1, t2, pet):
        """Flip-rate observations on the 64^3 grid."""
        self._obs = (np.ascontiguousarray(t1, np.uint8), np.ascontiguousarray(t2, np.uint8),
                     np.ascontiguousarray(pet, np.float32))
        self._c("gq_sur_obs_flip_upload", ptr(self._obs[0], C.c_uint8), ptr(self._obs[1], C.c_uint8),
                ptr(self._obs[2], C.c_float))

    def loglik_flip(self, params, u0, lik=(0.7, 0.25, 0.05, 0.05)):
        """(batch, 4): logL, L_T1, L_FLAIR, L_PET of the predictions."""
        p, u, _ = self._arr(params, u0)
        lk = np.ascontiguousarray(lik, np.float64)
        out = np.zeros((len(p), 4))
        for i in range(0, len(p), self.batch):
            b = min(self.batch, len(p) - i)
            self._c("gq_sur_loglik_flip", ptr(p[i:i + b], C.c_float), ptr(u[i:i + b], C.c_float), b,
                    ptr(lk, C.c_double), ptr(out[i:i + b], C.c_double))
        return out

    def timing(self):
        out = np.zeros(2)
        self._c("gq_sur_timing", ptr(out, C.c_double))
        return {"forward_ms": out[0], "backward_ms": out[1]}


def conv_flops(batch):
    """Multiply-add flops (x2) of the network's 3^3 convolutions for one
    forward pass, and for forward + backward (data and weight gradients)."""
    spec = [(8, 16, 64), (16, 32, 32), (32, 64, 16), (64, 128, 8), (192, 64, 16), (96, 32, 32), (48, 16, 64)]
    f = sum(2.0 * d ** 3 * 27 * ci * co for ci, co, d in spec) * batch
    # backward: data gradient (all but E1) + weight gradient
    b = sum(2.0 * d ** 3 * 27 * ci * co * (2 if i > 0 else 1) for i, (ci, co, d) in enumerate(spec)) * batch
    return f, f + b
