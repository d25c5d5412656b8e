"""Host-side mirror of the reference's block-operator API over the C-ABI.

Names follow the reference so that a caller of the MRAG/Glioma path finds the
same pieces (paths relative to the reference's source/ tree):

  Grid                              MRAGcore/MRAGrid.h:45-186 (+ device block data)
  Grid.getBlocksInfo                MRAGcore/MRAGrid.inl:191-219
  Grid.refine / Grid.compress       MRAGcore/MRAGrid.inl:248-355, :384-486
  Glioma_ReactionDiffusionOperator  Glioma/Glioma_ReactionDiffusionOperator.h:13-222
  UpdateTumor                       Glioma/Glioma_ReactionDiffusionOperator.h:227-266
  BlockProcessing.pipeline_process  MRAGmultithreading/MRAGBlockProcessing_TBB.h:476-512
  BlockProcessing.process           MRAGmultithreading/MRAGBlockProcessing_TBB.h:249-264
  BlockFWT.details                  MRAGcore/MRAGBlockFWT.h:448-478
  Science.AutomaticRefinement       MRAGscience/MRAGScienceCore.h:42-108
  Science.AutomaticCompression      MRAGscience/MRAGScienceCore.h:117-177
  Glioma_HG_UQ._reactionDiffusionStep / run / _dumpUQoutput
                                    Glioma/Glioma_HG_UQ.cpp:266-277, 511-588, 309-372
  HGG_Likelihood                    Likelihood/source/HGG_Likelihood.cpp:11-189

The RD functor and UpdateTumor execute as one fused kernel: pipeline_process
with the RD functor stages it, and the following process(UpdateTumor) runs
the step (a pipeline_process that is not followed by an UpdateTumor raises).
Every computation runs in libgq.so on the GPU; there is no CPU path here.
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import BlockInfo, FKParams, GQError, RunOpts, ptr

BS = 16


def _check(L, ctx, st):
    if st != _lib.GQ_OK:
        msg = L.gq_last_error(ctx).decode() if ctx else ""
        raise GQError(st, msg)


class Context:
    """A gq_ctx on one GPU (device memory + stream)."""

    def __init__(self, device=0):
        self.L = _lib.load()
        h = C.c_void_p()
        st = self.L.gq_ctx_create(C.byref(h), device)
        if st != _lib.GQ_OK:
            raise GQError(st, "gq_ctx_create(device=%d)" % device)
        self.h = h

    def close(self):
        if self.h:
            self.L.gq_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        _check(self.L, self.h, getattr(self.L, name)(self.h, *args))

    # a context holds one grid or ensemble at a time; creating another one
    # replaces it, and the older Python object must not address the new one
    owner = None

    def check_owner(self, obj):
        if self.owner is not obj:
            raise RuntimeError("this context now holds another grid/ensemble: create one context per grid")


class Grid:
    """Grid<W,B>(bpd,bpd,bpd) with Refiner_SpaceExtension(1, maxLevel) and
    Compressor(1); block data (phi, p_w, p_g) resident on the GPU."""

    FIELDS = {"phi": _lib.FIELD_PHI, "pw": _lib.FIELD_PWM, "pg": _lib.FIELD_PGM}

    def __init__(self, ctx, blocks_per_dim, max_level, wavelet_order=4):
        self.ctx = ctx
        self.bpd = blocks_per_dim
        self.max_level = max_level
        self.order = wavelet_order
        ctx.call("gq_grid_create", blocks_per_dim, max_level, wavelet_order)
        ctx.owner = self

    # -- topology -------------------------------------------------------
    def n_blocks(self):
        n = C.c_int()
        self.ctx.call("gq_grid_block_count", C.byref(n))
        return n.value

    def getBlocksInfo(self):
        n = self.n_blocks()
        arr = (BlockInfo * max(n, 1))()
        m = C.c_int()
        self.ctx.call("gq_grid_blocks_info", arr, n, C.byref(m))
        return [arr[i] for i in range(m.value)]

    def levels_indices(self):
        info = self.getBlocksInfo()
        lv = np.array([b.level for b in info], np.int32)
        ix = np.array([[b.index[0], b.index[1], b.index[2]] for b in info], np.int32).reshape(-1, 3)
        ids = np.array([b.block_id for b in info], np.int32)
        return lv, ix, ids

    def refine(self, block_ids):
        ids = np.ascontiguousarray(sorted(block_ids), np.int32)
        out = C.c_int()
        self.ctx.call("gq_refine_blocks", ptr(ids, C.c_int32), len(ids), C.byref(out))
        return out.value

    def compress(self, block_ids):
        ids = np.ascontiguousarray(sorted(block_ids), np.int32)
        out = C.c_int()
        self.ctx.call("gq_compress_blocks", ptr(ids, C.c_int32), len(ids), C.byref(out))
        return out.value

    # -- data -------------------------------------------------------------
    def write(self, name, blocks):
        a = np.ascontiguousarray(blocks, np.float32)
        n = self.n_blocks()
        if a.shape != (n, BS, BS, BS):
            raise ValueError("expected (%d,16,16,16), got %s" % (n, a.shape))
        self.ctx.call("gq_grid_write_field", self.FIELDS[name], ptr(a, C.c_float), n)

    def read(self, name):
        n = self.n_blocks()
        a = np.zeros((n, BS, BS, BS), np.float32)
        self.ctx.call("gq_grid_read_field", self.FIELDS[name], ptr(a, C.c_float), n)
        return a

    def sample(self, name, fn):
        """Fill a field from fn(x, y, z) evaluated at every block point
        (vertex-centred positions origin + h*i), arrays indexed [z,y,x]."""
        blocks = []
        for b in self.getBlocksInfo():
            r = np.arange(BS) * b.h
            z, y, x = np.meshgrid(b.origin[2] + r, b.origin[1] + r, b.origin[0] + r, indexing="ij")
            blocks.append(fn(x, y, z))
        self.write(name, np.stack(blocks))

    def write_volume(self, name, vol):
        """Point-sample a uniform volume vol[z,y,x] of gpd^3 = (bpd*16)^3
        points onto every leaf, on the device (gq_grid_write_volume)."""
        v = np.ascontiguousarray(vol, np.float32)
        self.ctx.call("gq_grid_write_volume", self.FIELDS[name], ptr(v, C.c_float), v.shape[0],
                      getattr(self.ctx, "sample", 0))

    def sample_volume(self, name, vol):
        """Point-sample a uniform volume vol[z,y,x] of gpd^3 = (bpd*16)^3 points
        onto every leaf (a leaf at level l takes every 2^(base-l)-th point)."""
        gpd = vol.shape[0]
        base = int(round(math.log2(gpd // BS)))
        blocks = []
        for b in self.getBlocksInfo():
            s = 1 << (base - b.level)
            o = [b.index[a] * BS * s for a in range(3)]
            blocks.append(vol[o[2]:o[2] + BS * s:s, o[1]:o[1] + BS * s:s, o[0]:o[0] + BS * s:s])
        self.write(name, np.stack(blocks))

    # -- uniform view -------------------------------------------------------
    def dump_uniform(self):
        """Glioma_HG_UQ::_dumpUQoutput: phi on the gpd^3 grid (float64, [z,y,x])."""
        gpd = self.bpd * BS
        out = np.zeros(gpd ** 3, np.float64)
        self.ctx.call("gq_dump_uniform", ptr(out, C.c_double), out.size)
        return out.reshape(gpd, gpd, gpd)


class Glioma_ReactionDiffusionOperator:
    """RD functor parameters (Dw, Dg, rho) in unit-cube units; stencil [-1,2)."""
    stencil_start = (-1, -1, -1)
    stencil_end = (2, 2, 2)

    def __init__(self, Dw, Dg, rho):
        self.Dw, self.Dg, self.rho = Dw, Dg, rho


class UpdateTumor:
    def __init__(self, dt):
        self.dt = dt


class BlockProcessing:
    """Drop-in for BlockProcessing_Pipeline_TBB on the GPU."""

    def __init__(self):
        self._staged = None

    def pipeline_process(self, vInfo, grid, boundaryInfo, op):
        if not isinstance(op, Glioma_ReactionDiffusionOperator):
            raise TypeError("only Glioma_ReactionDiffusionOperator runs on this path")
        self._staged = (grid, op)

    def process(self, vInfo, grid, op, granularity=4):
        if not isinstance(op, UpdateTumor):
            raise TypeError("only UpdateTumor runs on this path")
        if self._staged is None or self._staged[0] is not grid:
            raise RuntimeError("UpdateTumor without a staged Glioma_ReactionDiffusionOperator")
        rd = self._staged[1]
        self._staged = None
        fk_steps(grid, rd.Dw, rd.Dg, rd.rho, op.dt, 1)


def fk_steps(grid, Dw, Dg, rho, dt, n_steps):
    p = FKParams(Dw, Dg, rho, dt)
    grid.ctx.call("gq_fk_steps", C.byref(p), int(n_steps))


def step_timing(grid):
    ms = C.c_double()
    n = C.c_int()
    cells = C.c_int64()
    grid.ctx.call("gq_last_step_timing", C.byref(ms), C.byref(n), C.byref(cells))
    return ms.value, n.value, cells.value


def step_timing_drain(ctx):
    """(ms, cells) arrays of every step launch timed since the last drain
    (per-kernel timing) or of every chunk of steps (otherwise), in stream
    order (gq_step_timing_drain); waits for the timed work."""
    c = ctx.ctx if isinstance(ctx, Grid) else ctx
    n = C.c_int()
    c.call("gq_step_timing_drain", None, None, 0, C.byref(n))
    ms = np.zeros(max(n.value, 1), np.float32)
    cells = np.zeros(max(n.value, 1), np.int64)
    c.call("gq_step_timing_drain", ptr(ms, C.c_float), ptr(cells, C.c_int64), ms.size, C.byref(n))
    return ms[:n.value].astype(np.float64), cells[:n.value]


def set_step_timing(grid, per_kernel, stride=1):
    """per_kernel: time step-kernel launches with dispatch events, every
    `stride`-th launch (counted across calls); step_timing then reports the
    timed launches and their cell updates."""
    grid.ctx.call("gq_set_step_timing", int(stride) if per_kernel else 0)


def set_step_graphs(ctx, on):
    """gq_fk_steps as graph replays (default) or direct launches (same bits)."""
    ctx.call("gq_set_step_graphs", 1 if on else 0)


def set_autotune(ctx, mode=1, fixed_zt=16):
    """mode 1: autotune the step slab depth (16, 8, 4, 2 planes) with the
    cross-level face ghosts from a separate launch; 4: also the column kernel
    (depth 1) and every depth with the face ghosts evaluated in the step; 0:
    size rule; 2: fixed_zt (16/8/4/2/1), separate launch; 3: fixed_zt,
    evaluated in the step."""
    ctx.call("gq_set_autotune", int(mode), int(fixed_zt))


def set_step_coeffs(ctx, precomputed=True):
    """True (default): the step reads 1/K precomputed per cell (phi and 1/K,
    8 B per cell); False: it recomputes 1/K from pw, pg every step.  Same
    bits either way."""
    ctx.call("gq_set_step_coeffs", 1 if precomputed else 0)


def set_step_fusion(ctx, on=True):
    """True: pairs of steps over a work list without cross-level faces
    (uniform grids) run as one fused kernel launch; False (the context's
    default, measured faster): one launch per step.  Same bits either way."""
    ctx.call("gq_set_step_fusion", 1 if on else 0)


def set_step_persistent(ctx, mode=2):
    """Persistent chunks of steps (one cooperative launch per chunk between
    adaptations, grid-wide barriers between the face-ghost and step phases):
    0 never (the context's default, measured faster), 1 an autotuner
    candidate, 2 always.  Same bits in every mode."""
    ctx.call("gq_set_step_persistent", int(mode))


def persistent_info(ctx):
    """(last step call ran persistent chunks?, autotuner's persistent timing
    in us per step or 0)."""
    used = C.c_int()
    us = C.c_double()
    ctx.call("gq_persistent_info", C.byref(used), C.byref(us))
    return bool(used.value), us.value


def autotune_info(ctx):
    """(slab depth, face ghosts in the step?, {"<depth>": us, "<depth>f": us})
    of the last step call and the last autotune timings."""
    zt = C.c_int()
    fused = C.c_int()
    us = np.zeros(10)
    ctx.call("gq_autotune_info", C.byref(zt), C.byref(fused), ptr(us, C.c_double))
    names = ["%d" % d for d in (16, 8, 4, 2, 1)] + ["%df" % d for d in (16, 8, 4, 2, 1)]
    return zt.value, bool(fused.value), dict(zip(names, us.tolist()))


def host_stats(ctx):
    out = np.zeros(8)
    ctx.call("gq_host_stats", ptr(out, C.c_double))
    keys = ("tables_s", "plan_s", "fwt_s", "split_s", "collapse_s", "autotune_s", "n_tables", "n_fwt")
    return dict(zip(keys, out.tolist()))


def graph_info(ctx):
    """(captures, relaunches, cached graphs) of the context's step chunks."""
    c = C.c_int64()
    r = C.c_int64()
    k = C.c_int()
    ctx.call("gq_graph_info", C.byref(c), C.byref(r), C.byref(k))
    return c.value, r.value, k.value


class BlockFWT:
    @staticmethod
    def details(grid):
        n = grid.n_blocks()
        out = np.zeros(n, np.float64)
        grid.ctx.call("gq_fwt_details", ptr(out, C.c_double), n)
        return out


class Science:
    @staticmethod
    def AutomaticRefinement(grid, tol, max_level=-1, max_loops=-1):
        out = C.c_int()
        grid.ctx.call("gq_auto_refine", tol, max_level, max_loops, C.byref(out))
        return out.value

    @staticmethod
    def AutomaticCompression(grid, tol, max_loops=-1):
        out = C.c_int()
        grid.ctx.call("gq_auto_compress", tol, max_loops, C.byref(out))
        return out.value


def n_steps_until(T, dt):
    """Number of iterations of `while (t <= tend) { step; t += dt; }`
    (Glioma_HG_UQ.cpp:544-552) with t accumulated in double."""
    t = 0.0
    n = 0
    while t <= T:
        n += 1
        t += dt
    return n


class Glioma_HG_UQ:
    """The forward-solve driver of Glioma_HG_UQ::run (Glioma_HG_UQ.cpp:511-588).

    Dw_cm2_day / rho_day / T_days as in HGG_InputParameters.txt; L is the
    physical edge of the unit cube (cm).  Adaptation runs every
    `adapt_every` steps (integer schedule) or, when `dumpfreq` is given, with
    the reference's `t >= whenToWrite` schedule in days."""

    def __init__(self, grid, L_cm, cfl=0.99, adaptive=True, tol_refine=1e-4, tol_compress=1e-5,
                 adapt_every=None, dumpfreq=None):
        self.grid = grid
        self.L = L_cm
        self.cfl = cfl
        self.adaptive = adaptive
        self.tol_r = tol_refine
        self.tol_c = tol_compress
        self.adapt_every = adapt_every
        self.dumpfreq = dumpfreq
        self.bp = BlockProcessing()
        self.history = []

    def params(self, Dw_cm2_day, rho_day):
        """(Dw, Dg, rho, dt) as Glioma_HG_UQ::run forms them (:519-539): the
        inputs are read into Real = float, Dw/L^2 and 0.1*Dw are evaluated in
        double and stored as float, dt is formed in double from those floats."""
        f32 = np.float32
        Dw = float(f32(float(f32(Dw_cm2_day)) / (self.L * self.L)))
        Dg = float(f32(0.1 * Dw))
        rho = float(f32(rho_day))
        h = 1.0 / (BS * self.grid.bpd)
        dt = self.cfl * h * h / (2.0 * 3 * max(Dw, Dg))
        return Dw, Dg, rho, dt

    def _adapt(self):
        g = self.grid
        nr = Science.AutomaticRefinement(g, self.tol_r, g.max_level, 1)
        nc = Science.AutomaticCompression(g, self.tol_c, -1)
        self.history.append((nr, nc, g.n_blocks()))

    def run(self, Dw_cm2_day, rho_day, T_days, chunk_cb=None):
        Dw, Dg, rho, dt = self.params(Dw_cm2_day, rho_day)
        n_total = n_steps_until(T_days, dt)
        done = 0
        if not self.adaptive:
            fk_steps(self.grid, Dw, Dg, rho, dt, n_total)
            return n_total, dt
        if self.dumpfreq is not None:
            t = 0.0
            when = self.dumpfreq
            while t <= T_days:
                fk_steps(self.grid, Dw, Dg, rho, dt, 1)
                t += dt
                done += 1
                if t >= when:
                    self._adapt()
                    when += self.dumpfreq
        else:
            k = self.adapt_every
            while done < n_total:
                m = min(k, n_total - done)
                fk_steps(self.grid, Dw, Dg, rho, dt, m)
                done += m
                if chunk_cb:
                    chunk_cb(done)
                if m == k:
                    self._adapt()
        # final refinement before the dump (:574-575)
        Science.AutomaticRefinement(self.grid, self.tol_r, self.grid.max_level, 1)
        return n_total, dt


class HGG_Likelihood:
    """Reference likelihood model on the device (HGG_Likelihood.cpp:11-189).

    params = (PETsigma2, PETscale, T1uc, T2uc, slope) as LikelihoodInput.txt;
    flags as -selectivePoints (default 1), -stepPet (3), -stepTi (2)."""

    def __init__(self, grid, pet, t1, t2, params, points_xyz=None, selective=True, step_pet=3, step_ti=2):
        self.grid = grid
        gpd = grid.bpd * BS
        self.params = np.ascontiguousarray(params, np.float64)
        self.selective = selective
        self.step_pet = step_pet
        self.step_ti = step_ti
        pts = np.zeros((0, 3), np.int32) if points_xyz is None else np.ascontiguousarray(points_xyz, np.int32)
        self._keep = [np.ascontiguousarray(a, np.float64) for a in (pet, t1, t2)] + [pts]
        if any(a.shape != (gpd, gpd, gpd) for a in self._keep[:3]):
            raise ValueError("PET, T1, T2 must be volumes of the model's %d^3 grid" % gpd)
        if pts.ndim != 2 or pts.shape[1] != 3:
            raise ValueError("points must be an (n, 3) array")
        grid.ctx.call("gq_obs_ref_upload", gpd, ptr(self._keep[0], C.c_double), ptr(self._keep[1], C.c_double),
                      ptr(self._keep[2], C.c_double), ptr(pts, C.c_int32), len(pts))

    def run(self):
        out = np.zeros(4)
        self.grid.ctx.call("gq_loglik_ref", ptr(self.params, C.c_double), int(self.selective), self.step_pet,
                           self.step_ti, ptr(out, C.c_double))
        return out  # logL, L_pet, L_T1, L_T2


class _SampleCtx:
    """Context proxy that addresses one sample of an ensemble."""

    def __init__(self, ctx, sample):
        self.base = ctx
        self.sample = sample
        self.L = ctx.L
        self.h = ctx.h

    def call(self, name, *args):
        self.base.call("gq_sample_select", self.sample)
        self.base.call(name, *args)


class Ensemble:
    """nsamples independent grids of one configuration advancing together
    (gq_ensemble_*): one step-kernel launch per step covers the blocks of
    every running sample; each sample's result equals a one-sample run."""

    def __init__(self, ctx, nsamples, blocks_per_dim, max_level, wavelet_order=4):
        self.ctx = ctx
        self.n = nsamples
        self.bpd = blocks_per_dim
        self.max_level = max_level
        self.order = wavelet_order
        ctx.call("gq_ensemble_create", nsamples, blocks_per_dim, max_level, wavelet_order)
        ctx.owner = self

    def reset(self):
        """Every sample back to its uniform base-level grid, keeping the device
        pool (gq_ensemble_reset); fields must be written again."""
        self.ctx.call("gq_ensemble_reset")

    def sample(self, s):
        """Grid view of sample s (Grid methods address that sample)."""
        if not 0 <= s < self.n:
            raise IndexError(s)
        g = Grid.__new__(Grid)
        g.ctx = _SampleCtx(self.ctx, s)
        g.bpd, g.max_level, g.order = self.bpd, self.max_level, self.order
        return g

    def write_volume(self, name, vol, sample=-1):
        v = np.ascontiguousarray(vol, np.float32)
        self.ctx.call("gq_grid_write_volume", Grid.FIELDS[name], ptr(v, C.c_float), v.shape[0], int(sample))

    def block_counts(self):
        out = np.zeros(self.n, np.int32)
        self.ctx.call("gq_ensemble_block_counts", ptr(out, C.c_int32))
        return out

    def run(self, params, n_steps, adapt_every=10, max_level=None, tol_refine=1e-3, tol_compress=1e-4, graph=True):
        """params: (nsamples, 4) array of (Dw, Dg, rho, dt) in unit-cube units;
        n_steps: steps per sample."""
        prm = np.ascontiguousarray(params, np.float64).reshape(self.n, 4)
        ns = np.ascontiguousarray(n_steps, np.int32).reshape(self.n)
        arr = (FKParams * self.n)(*[FKParams(*map(float, r)) for r in prm])
        o = RunOpts(int(adapt_every), int(self.max_level if max_level is None else max_level), float(tol_refine),
                    float(tol_compress), _lib.RUN_GRAPH if graph else 0)
        self.ctx.call("gq_ensemble_run", arr, ptr(ns, C.c_int32), C.byref(o))
        return self.stats()

    def stats(self):
        out = np.zeros(8)
        self.ctx.call("gq_ensemble_stats", ptr(out, C.c_double))
        keys = ("cell_updates", "device_ms", "chunks", "adaptations", "adapt_s", "final_blocks", "zt", "graph")
        return dict(zip(keys, out.tolist()))

    def set_observations(self, t1, t2, pet):
        """Flip-rate model observations on the gpd^3 grid or on a grid coarser
        by an integer factor (gq_obs_flip_upload)."""
        keep = (np.ascontiguousarray(t1, np.uint8), np.ascontiguousarray(t2, np.uint8),
                np.ascontiguousarray(pet, np.float32))
        gpd = keep[0].shape[0]
        if any(a.shape != (gpd, gpd, gpd) for a in keep):
            raise ValueError("observations must be three cubic volumes of one size")
        self.ctx.call("gq_obs_flip_upload", gpd, ptr(keep[0], C.c_uint8), ptr(keep[1], C.c_uint8),
                      ptr(keep[2], C.c_float))

    def loglik_flip(self, t1=None, t2=None, pet=None, th1=0.7, th2=0.25, flip=0.05, sigma=0.05):
        if t1 is not None:
            self.set_observations(t1, t2, pet)
        prm = np.array([th1, th2, flip, sigma], np.float64)
        out = np.zeros((self.n, 4))
        self.ctx.call("gq_ensemble_loglik_flip", ptr(prm, C.c_double), ptr(out, C.c_double))
        return out  # per sample: logL, L_T1, L_FLAIR, L_PET

    def ic_gaussian(self, ic):
        """Gaussian initial condition per sample: ic (nsamples, 5) = x0, y0, z0,
        amplitude, sigma in voxel units (uniform grids)."""
        a = np.ascontiguousarray(ic, np.float64).reshape(self.n, 5)
        self.ctx.call("gq_ensemble_ic_gaussian", ptr(a, C.c_double))

    def adjoint_flip(self, params, n_steps, ckpt_every=50, ic=None, th1=0.7, th2=0.25, flip=0.05, sigma=0.05):
        """Forward + discrete adjoint (gq_adjoint_flip).  Returns (loglik
        (nsamples, 4), grad (nsamples, 6) = dL/d(log Dw, log rho, x0, y0, z0, A))."""
        prm = np.ascontiguousarray(params, np.float64).reshape(self.n, 4)
        arr = (FKParams * self.n)(*[FKParams(*map(float, r)) for r in prm])
        lik = np.array([th1, th2, flip, sigma], np.float64)
        ll = np.zeros((self.n, 4))
        gr = np.zeros((self.n, 6))
        icp = None
        if ic is not None:
            icp = np.ascontiguousarray(ic, np.float64).reshape(self.n, 5)
        self.ctx.call("gq_adjoint_flip", arr, int(n_steps), int(ckpt_every), ptr(lik, C.c_double),
                      None if icp is None else ptr(icp, C.c_double), ptr(ll, C.c_double), ptr(gr, C.c_double))
        return ll, gr


class FlipLikelihood:
    """Flip-rate Bernoulli (T1Gd, FLAIR) + Gaussian PET likelihood on the device."""

    def __init__(self, grid, t1, t2, pet, th1=0.7, th2=0.25, flip=0.05, sigma=0.05):
        """t1, t2, pet: observation volumes on the model's gpd^3 grid or on a
        grid coarser by an integer factor (gq_obs_flip_upload)."""
        self.grid = grid
        self._keep = (np.ascontiguousarray(t1, np.uint8), np.ascontiguousarray(t2, np.uint8),
                      np.ascontiguousarray(pet, np.float32))
        gpd = self._keep[0].shape[0]
        if any(a.shape != (gpd, gpd, gpd) for a in self._keep):
            raise ValueError("observations must be three cubic volumes of one size")
        self.params = np.array([th1, th2, flip, sigma], np.float64)
        grid.ctx.call("gq_obs_flip_upload", gpd, ptr(self._keep[0], C.c_uint8), ptr(self._keep[1], C.c_uint8),
                      ptr(self._keep[2], C.c_float))

    def loglik(self):
        out = np.zeros(4)
        self.grid.ctx.call("gq_loglik_flip", ptr(self.params, C.c_double), ptr(out, C.c_double))
        return out  # logL, L_T1, L_FLAIR, L_PET


class Topology:
    """Host-only Grid topology (gq_topo_*): refine/compress planning and ghost
    expansion without block data and without a GPU."""

    def __init__(self, blocks_per_dim, max_level, wavelet_order=4):
        self.L = _lib.load()
        h = C.c_void_p()
        st = self.L.gq_topo_create(C.byref(h), blocks_per_dim, max_level, wavelet_order)
        if st != _lib.GQ_OK:
            raise GQError(st, "gq_topo_create")
        self.h = h

    def __del__(self):
        try:
            if self.h:
                self.L.gq_topo_destroy(self.h)
        except Exception:
            pass

    def _c(self, st):
        if st != _lib.GQ_OK:
            raise GQError(st, "topology call failed")

    def leaves(self):
        n = C.c_int()
        self._c(self.L.gq_topo_leaves(self.h, None, None, None, 0, C.byref(n)))
        m = n.value
        lv = np.zeros(m, np.int32)
        ix = np.zeros((m, 3), np.int32)
        ids = np.zeros(m, np.int32)
        self._c(self.L.gq_topo_leaves(self.h, ptr(lv, C.c_int32), ptr(ix, C.c_int32), ptr(ids, C.c_int32), m,
                                      C.byref(n)))
        return lv, ix, ids

    def refine(self, ids):
        a = np.ascontiguousarray(sorted(ids), np.int32)
        out = C.c_int()
        self._c(self.L.gq_topo_refine(self.h, ptr(a, C.c_int32), len(a), C.byref(out)))
        return out.value

    def compress(self, ids):
        a = np.ascontiguousarray(sorted(ids), np.int32)
        out = C.c_int()
        self._c(self.L.gq_topo_compress(self.h, ptr(a, C.c_int32), len(a), C.byref(out)))
        return out.value

    def point_terms(self, leaf, loc, cap=4096):
        lo = np.ascontiguousarray(loc, np.int32)
        ol = np.zeros(cap, np.int32)
        oc = np.zeros(cap, np.int32)
        ow = np.zeros(cap, np.float64)
        n = C.c_int()
        self._c(self.L.gq_topo_point_terms(self.h, leaf, ptr(lo, C.c_int32), cap, ptr(ol, C.c_int32),
                                           ptr(oc, C.c_int32), ptr(ow, C.c_double), C.byref(n)))
        m = n.value
        return ol[:m].copy(), oc[:m].copy(), ow[:m].copy()
