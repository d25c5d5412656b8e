"""ctypes bindings of the CPU oracle (fk_oracle.c) -- TEST INFRASTRUCTURE ONLY.

Every function restates a reference function (file:line cited in fk_oracle.c).
Loaded from oracle/liboracle.so, built by `make -C oracle`.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        u8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        L.fk_block_f64.argtypes = [dp, dp, dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.fk_block_f32ref.argtypes = [fp, fp, fp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, fp]
        L.fk_update_f64.argtypes = [dp, dp, C.c_long, C.c_double]
        L.fk_update_f32ref.argtypes = [fp, fp, C.c_long, C.c_double]
        L.fk_uniform_steps_f64.argtypes = [dp, dp, dp, C.c_long, C.c_int, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_double, C.c_int, C.c_int]
        L.fk_uniform_steps_f32ref.argtypes = [fp, fp, fp, C.c_long, C.c_int, C.c_float, C.c_float, C.c_float,
                                              C.c_float, C.c_double, C.c_int, C.c_int]
        L.og_point_terms.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip, ip, C.c_int, C.c_int, ip, C.c_int,
                                     ip, ip, dp]
        L.og_fill_lab_f64.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip, ip, C.c_int, dp, C.c_int, C.c_int,
                                      C.c_int, dp]
        L.fk_adaptive_steps_f64.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip, ip, C.c_int, dp, dp, dp,
                                            C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
        L.fwt_block_detail_f64.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.fwt_block_detail_f64.restype = C.c_double
        L.dump_uniform_f64.argtypes = [C.c_int, C.c_int, ip, ip, dp, C.c_int, dp]
        L.lik_ref_f64.argtypes = [dp, dp, dp, dp, C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int, C.c_int,
                                  dp, C.c_int, C.POINTER(C.c_longdouble)]
        L.lik_ref_f64.restype = C.c_longdouble
        L.lik_flip_f64.argtypes = [dp, u8, u8, fp, C.c_long, C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.lik_flip_f64.restype = C.c_double
        L.adj_fk_f64.argtypes = [dp, dp, C.c_long, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_int, dp, dp, dp, dp, dp]
        L.adj_fk_f64_ckpt.argtypes = [dp, dp, C.c_long, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_int, C.c_int, C.c_int, dp, C.c_void_p, C.c_void_p, C.c_double, dp, dp, dp]
        _LIB = L
    return _LIB


def fk_block(lab_phi, lab_pw, lab_pg, h, Dw, Dg, rho, mode="f64"):
    """RD functor on one ghosted lab (halo 1). Returns dphidt (B^3)."""
    L = lab_phi.shape[0]
    B = L - 2
    if mode == "f64":
        a = [np.ascontiguousarray(x, np.float64) for x in (lab_phi, lab_pw, lab_pg)]
        out = np.zeros((B, B, B), np.float64)
        lib().fk_block_f64(a[0], a[1], a[2], B, h, Dw, Dg, rho, out)
    else:
        a = [np.ascontiguousarray(x, np.float32) for x in (lab_phi, lab_pw, lab_pg)]
        out = np.zeros((B, B, B), np.float32)
        lib().fk_block_f32ref(a[0], a[1], a[2], B, h, Dw, Dg, rho, out)
    return out


def fk_update(phi, dphidt, dt, mode="f64"):
    if mode == "f64":
        p = np.ascontiguousarray(phi, np.float64).copy()
        lib().fk_update_f64(p, np.ascontiguousarray(dphidt, np.float64), p.size, dt)
    else:
        p = np.ascontiguousarray(phi, np.float32).copy()
        lib().fk_update_f32ref(p, np.ascontiguousarray(dphidt, np.float32), p.size, dt)
    return p


def fk_uniform_steps(phi, pw, pg, B, h, Dw, Dg, rho, dt, n_steps, mode="f64", threads=0):
    """n_steps of the uniform periodic block grid; volumes (N,N,N) z,y,x order."""
    N = phi.shape[0]
    if mode == "f64":
        p = np.ascontiguousarray(phi, np.float64).copy()
        rc = lib().fk_uniform_steps_f64(p, np.ascontiguousarray(pw, np.float64), np.ascontiguousarray(pg, np.float64),
                                        N, B, h, Dw, Dg, rho, dt, n_steps, threads)
    else:
        p = np.ascontiguousarray(phi, np.float32).copy()
        rc = lib().fk_uniform_steps_f32ref(p, np.ascontiguousarray(pw, np.float32),
                                           np.ascontiguousarray(pg, np.float32), N, B, h, Dw, Dg, rho, dt,
                                           n_steps, threads)
    if rc != 0:
        raise RuntimeError("oracle uniform steps failed")
    return p


class LeafGrid:
    """Oracle-side description of a leaf-block hierarchy (levels, indices)."""

    def __init__(self, levels, indices, B, max_level, order):
        self.level = np.ascontiguousarray(levels, np.int32)
        self.index = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
        self.B = B
        self.max_level = max_level
        self.order = order
        nF = 1 << max_level
        lut = -np.ones((nF, nF, nF), np.int32)  # [z][y][x]
        for i, (l, ix) in enumerate(zip(self.level, self.index)):
            s = 1 << (max_level - l)
            x, y, z = ix * s
            if (lut[z:z + s, y:y + s, x:x + s] >= 0).any():
                raise ValueError("overlapping leaves")
            lut[z:z + s, y:y + s, x:x + s] = i
        if (lut < 0).any():
            raise ValueError("leaves do not cover the domain")
        self.lut = np.ascontiguousarray(lut.reshape(-1))

    @property
    def n(self):
        return len(self.level)

    def point_terms(self, leaf, loc, cap=4096):
        ol = np.zeros(cap, np.int32)
        oc = np.zeros(cap, np.int32)
        ow = np.zeros(cap, np.float64)
        m = lib().og_point_terms(self.n, self.max_level, self.B, self.level, self.index.reshape(-1), self.lut,
                                 self.order, leaf, np.asarray(loc, np.int32), cap, ol, oc, ow)
        if m < 0:
            raise RuntimeError("ghost expansion overflow")
        return ol[:m].copy(), oc[:m].copy(), ow[:m].copy()

    def fill_lab(self, data, leaf, s0, s1):
        L = self.B + s1 - s0 - 1
        lab = np.zeros((L, L, L), np.float64)
        rc = lib().og_fill_lab_f64(self.n, self.max_level, self.B, self.level, self.index.reshape(-1), self.lut,
                                   self.order, np.ascontiguousarray(data, np.float64), leaf, s0, s1, lab)
        if rc != 0:
            raise RuntimeError("ghost expansion overflow")
        return lab

    def fk_steps(self, phi, pw, pg, Dw, Dg, rho, dt, n_steps, threads=0):
        p = np.ascontiguousarray(phi, np.float64).copy()
        rc = lib().fk_adaptive_steps_f64(self.n, self.max_level, self.B, self.level, self.index.reshape(-1),
                                         self.lut, self.order, p, np.ascontiguousarray(pw, np.float64),
                                         np.ascontiguousarray(pg, np.float64), Dw, Dg, rho, dt, n_steps, threads)
        if rc != 0:
            raise RuntimeError("ghost expansion overflow")
        return p

    def fwt_details(self, phi, which=None, threads=8):
        """Per-leaf detail magnitude (FWTReport::getOverAll_DetailMaxMag) of
        the leaves `which` (default all; others are left 0).  The C calls
        release the GIL, so leaves run on `threads` host threads."""
        s0, s1 = (-2, 4) if self.order == 4 else (-1, 2)
        out = np.zeros(self.n)
        L = self.B + s1 - s0 - 1
        data = np.ascontiguousarray(phi, np.float64)
        todo = range(self.n) if which is None else list(which)

        def one(b):
            lab = self.fill_lab(data, b, s0, s1)
            out[b] = lib().fwt_block_detail_f64(lab, self.B, s0, L, self.order)

        if threads > 1 and len(todo) > 8:
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(threads) as ex:
                list(ex.map(one, todo))
        else:
            for b in todo:
                one(b)
        return out

    def dump_uniform(self, phi, gpd):
        out = np.zeros(gpd ** 3, np.float64)
        lib().dump_uniform_f64(self.n, self.B, self.level, self.index.reshape(-1),
                               np.ascontiguousarray(phi, np.float64), gpd, out)
        return out.reshape(gpd, gpd, gpd)


def lik_ref(model, pet, t1, t2, params, selective=True, step_pet=3, step_ti=2, points=None):
    """HGG_Likelihood::run (returns log L and its three parts)."""
    nz, ny, nx = model.shape
    if points is None:
        points = np.zeros((3, 1))
        npts = 0
    else:
        npts = points.shape[1] if points.ndim == 2 else 0
    pts = np.ascontiguousarray(points, np.float64)
    parts = (C.c_longdouble * 3)()
    v = lib().lik_ref_f64(*(np.ascontiguousarray(a, np.float64) for a in (model, pet, t1, t2)), nx, ny, nz,
                          np.ascontiguousarray(params, np.float64), int(selective), step_pet, step_ti, pts, npts,
                          parts)
    return float(v), [float(parts[i]) for i in range(3)]


def lik_flip(u, t1, t2, pet, th1=0.7, th2=0.25, eps=0.05, sigma=0.05):
    parts = np.zeros(3)
    v = lib().lik_flip_f64(np.ascontiguousarray(u, np.float64).reshape(-1),
                           np.ascontiguousarray(t1, np.uint8).reshape(-1),
                           np.ascontiguousarray(t2, np.uint8).reshape(-1),
                           np.ascontiguousarray(pet, np.float32).reshape(-1), u.size, th1, th2, eps, sigma, parts)
    return float(v), parts


def adjoint(u0, pw, pg, h, Dw, Dg, rho, dt, n_steps, dLduN_fn):
    """FP64 forward + discrete adjoint (adjoint_oracle.c) on a uniform periodic
    grid.  dLduN_fn(uN) -> cotangent at the final state.  Returns (uN, dL/du0,
    dL/dlogDw, dL/dlogrho); the forward is run twice (once to get uN)."""
    N = u0.shape[0]
    a = [np.ascontiguousarray(x, np.float64) for x in (pw, pg, u0)]
    uN = np.zeros_like(a[2])
    d0 = np.zeros_like(a[2])
    g = np.zeros(2)
    zero = np.zeros_like(a[2])
    rc = lib().adj_fk_f64(a[0], a[1], N, h, Dw, Dg, rho, dt, n_steps, a[2], zero, uN, d0, g)
    if rc != 0:
        raise MemoryError("adjoint oracle allocation failed")
    ct = np.ascontiguousarray(dLduN_fn(uN), np.float64)
    rc = lib().adj_fk_f64(a[0], a[1], N, h, Dw, Dg, rho, dt, n_steps, a[2], ct, uN, d0, g)
    if rc != 0:
        raise MemoryError("adjoint oracle allocation failed")
    return uN, d0, g[0], g[1]


def adjoint_ckpt(u0, pw, pg, h, Dw, Dg, rho, dt, n_steps, dLduN_fn=None, ckpt_every=50, threads=0, pet=None,
                 sigma=0.05):
    """adjoint() with checkpoints every ckpt_every steps and recomputation
    (memory of about n_steps/ckpt_every + ckpt_every states), OpenMP.  The
    terminal cotangent is dLduN_fn(uN), or, with pet given (NaN where not
    observed), that of the Gaussian PET term in one forward pass."""
    N = u0.shape[0]
    a = [np.ascontiguousarray(x, np.float64) for x in (pw, pg, u0)]
    uN = np.zeros_like(a[2])
    d0 = np.zeros_like(a[2])
    g = np.zeros(2)
    if pet is not None:
        pf = np.ascontiguousarray(pet, np.float32)
        rc = lib().adj_fk_f64_ckpt(a[0], a[1], N, h, Dw, Dg, rho, dt, n_steps, ckpt_every, threads, a[2], None,
                                   pf.ctypes.data_as(C.c_void_p), sigma, uN, d0, g)
    else:
        rc = lib().adj_fk_f64_ckpt(a[0], a[1], N, h, Dw, Dg, rho, dt, n_steps, ckpt_every, threads, a[2], None,
                                   None, sigma, uN, d0, g)
        if rc == 0:
            ct = np.ascontiguousarray(dLduN_fn(uN), np.float64)
            rc = lib().adj_fk_f64_ckpt(a[0], a[1], N, h, Dw, Dg, rho, dt, n_steps, ckpt_every, threads, a[2],
                                       ct.ctypes.data_as(C.c_void_p), None, sigma, uN, d0, g)
    if rc != 0:
        raise MemoryError("adjoint oracle allocation failed")
    return uN, d0, g[0], g[1]


def _pdisp(i, c, n):
    d = i - c
    return d - n * np.floor(d / n + 0.5)


def ic_gaussian(n, x0, y0, z0, amp, sigma):
    """Gaussian initial condition A exp(-|x - x0|^2 / (2 sigma^2)) on the n^3
    voxel grid (periodic signed displacement), FP64 [z, y, x]."""
    i = np.arange(n, dtype=np.float64)
    dz, dy, dx = np.meshgrid(_pdisp(i, z0, n), _pdisp(i, y0, n), _pdisp(i, x0, n), indexing="ij")
    return amp * np.exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * sigma * sigma))


def ic_grad(dLdu0, x0, y0, z0, amp, sigma):
    """Chain rule of ic_gaussian: dL/d(x0, y0, z0, A) from dL/du0."""
    n = dLdu0.shape[0]
    i = np.arange(n, dtype=np.float64)
    dz, dy, dx = np.meshgrid(_pdisp(i, z0, n), _pdisp(i, y0, n), _pdisp(i, x0, n), indexing="ij")
    e = np.exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * sigma * sigma))
    lu = dLdu0 * amp * e / (sigma * sigma)
    return np.array([np.sum(lu * dx), np.sum(lu * dy), np.sum(lu * dz), np.sum(dLdu0 * e)])
