"""GPU parity of the FK forward solve against the FP64 CPU oracle.

Bar (FP32 field vs the reference's FP64 arithmetic restated by the oracle):
relative L2 <= 1e-5 and max-abs <= 1e-5 on the field; the cells zeroed by
the reference's CSF quirk (0/0 -> NaN -> clamp to 0) are zero in both.
"""
import numpy as np
import pytest

import gliomauq_amd as gq
from gliomauq_amd import anatomy as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L_CM = 17.0


def rel_l2(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def zero_mismatch(a, b):
    return int(np.count_nonzero(((a == 0) != (b == 0)) & (np.maximum(a, b) > 1e-30)))


def uniform_case(n, seed=0):
    pwm, pgm, u0, c = A.config_inputs(n, L_CM, seed)
    return pwm, pgm, u0


@pytest.mark.parametrize("n,steps,order", [(64, 60, 2), (128, 40, 4)])
def test_uniform_grid_parity(ctx, n, steps, order):
    pwm, pgm, u0 = uniform_case(n, seed=1)
    # a larger initial tumour so the front crosses many cells
    u0 = np.maximum(u0, A.initial_condition(n, (n // 2, n // 2, n // 2 + 4), 0.6, 3.0))
    bpd = n // 16
    lvl = int(np.log2(bpd))
    g = gq.Grid(ctx, bpd, lvl, order)
    for name, v in (("phi", u0), ("pw", pwm), ("pg", pgm)):
        g.sample_volume(name, v)
    Dw = 0.0013 / L_CM ** 2 * 40  # faster diffusion so 40-60 steps move the front
    Dg, rho = 0.1 * Dw, 0.025 * 20
    h = 1.0 / n
    dt = 0.9 * h * h / (6 * Dw)
    gq.fk_steps(g, Dw, Dg, rho, dt, steps)
    u_gpu = g.dump_uniform()
    u_ref = O.fk_uniform_steps(u0.astype(np.float64), pwm.astype(np.float64), pgm.astype(np.float64), 16, h, Dw,
                               Dg, rho, dt, steps)
    e2 = rel_l2(u_gpu, u_ref)
    einf = float(np.max(np.abs(u_gpu - u_ref)))
    print("uniform n=%d steps=%d: rel L2 %.3e, Linf %.3e" % (n, steps, e2, einf))
    assert e2 <= 1e-5 and einf <= 1e-5
    assert zero_mismatch(u_gpu, u_ref) == 0
    assert (u_ref > 0).sum() > 1000


def test_pure_reaction_matches_recurrence(ctx):
    n = 32
    rng = np.random.default_rng(2)
    phi = (0.5 * rng.random((n, n, n))).astype(np.float32)
    pw = rng.random((n, n, n)).astype(np.float32)
    pg = (1 - pw).astype(np.float32)
    g = gq.Grid(ctx, 2, 1, 2)
    for name, v in (("phi", phi), ("pw", pw), ("pg", pg)):
        g.sample_volume(name, v)
    dt, rho = 0.8, 0.05
    gq.fk_steps(g, 0.0, 0.0, rho, dt, 10)
    u = g.dump_uniform()
    e = phi.astype(np.float64)
    for _ in range(10):
        e = np.clip(e + dt * rho * e * (1 - e), 0, 1)
    assert np.max(np.abs(u - e)) < 2e-6


def test_csf_quirk_cells_are_zero(ctx):
    n = 32
    phi = np.full((n, n, n), 0.3, np.float32)
    pw = np.full((n, n, n), 0.7, np.float32)
    pg = np.full((n, n, n), 0.3, np.float32)
    pw[5, 6, 7] = pg[5, 6, 7] = 0
    pw[16, 15, 0] = pg[16, 15, 0] = 0  # on a block face (x = 0 wraps to x = 31)
    g = gq.Grid(ctx, 2, 1, 2)
    for name, v in (("phi", phi), ("pw", pw), ("pg", pg)):
        g.sample_volume(name, v)
    gq.fk_steps(g, 1e-6, 1e-7, 0.02, 1.0, 1)
    u = g.dump_uniform()
    zeros = set(map(tuple, np.argwhere(u == 0)))
    exp = set()
    for c in [(5, 6, 7), (16, 15, 0)]:
        exp.add(c)
        for d in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
            exp.add(tuple((np.array(c) + d) % n))
    assert zeros == exp


def test_fk_steps_rejects_bad_parameters(ctx):
    g = gq.Grid(ctx, 2, 1, 2)
    with pytest.raises(gq._lib.GQError) as e:
        gq.fk_steps(g, float("nan"), 0.0, 0.1, 0.5, 1)
    assert e.value.status == 7
    with pytest.raises(gq._lib.GQError) as e:
        gq.fk_steps(g, 1e-6, 1e-7, 0.1, -1.0, 1)
    assert e.value.status == 7
    with pytest.raises(ValueError):
        g.write("phi", np.zeros((3, 16, 16, 16), np.float32))
    a = np.zeros((3, 16, 16, 16), np.float32)  # the C-ABI rejects a wrong block count itself
    with pytest.raises(gq._lib.GQError) as e:
        g.ctx.call("gq_grid_write_field", 0, gq._lib.ptr(a, gq._lib.C.c_float), 3)
    assert e.value.status == 1


def test_step_timing_every_kth_launch(ctx):
    """gq_set_step_timing(k): events on every k-th step launch, counted
    across calls; the field is the untimed run's to the bit."""
    n = 32
    rng = np.random.default_rng(5)
    phi = (0.4 * rng.random((n, n, n))).astype(np.float32)
    pw = rng.random((n, n, n)).astype(np.float32)
    pg = (1 - pw).astype(np.float32)
    outs = []
    for stride in (0, 1, 3):
        g = gq.Grid(ctx, 2, 1, 2)
        for name, v in (("phi", phi), ("pw", pw), ("pg", pg)):
            g.sample_volume(name, v)
        gq.set_step_timing(g, stride > 0, max(stride, 1))
        gq.step_timing_drain(g)  # entries of earlier work on this context
        got = []
        for m in (10, 10, 4):
            gq.fk_steps(g, 1e-5, 1e-6, 0.03, 0.5, m)
            got.append(gq.step_timing(g))
        outs.append(g.dump_uniform())
        cells = 8 * 4096
        if stride == 1:
            assert [k for _, k, _ in got] == [10, 10, 4]
        elif stride == 3:
            # launches 0..23: timed at 0, 3, ..., 21 -> 4 in the first call (0, 3, 6, 9),
            # 3 in the second (12, 15, 18), 1 in the third (21)
            assert [k for _, k, _ in got] == [4, 3, 1]
        if stride > 0:
            for ms, k, c in got:
                assert c == k * cells and ms > 0
            # every timed launch of the three calls, in order, then nothing
            n = gq._lib.C.c_int()
            ms1 = np.zeros(1, np.float32)
            if sum(k for _, k, _ in got) > 1:  # a capacity below the count is an error, entries kept
                with pytest.raises(gq._lib.GQError) as e:
                    g.ctx.call("gq_step_timing_drain", gq._lib.ptr(ms1, gq._lib.C.c_float), None, 1,
                               gq._lib.C.byref(n))
                assert e.value.status == 1
            ms, cl = gq.step_timing_drain(g)
            assert ms.size == sum(k for _, k, _ in got) and np.all(ms > 0) and np.all(cl == cells)
            assert abs(ms.sum() - sum(t for t, _, _ in got)) <= 1e-3 * ms.sum()
            assert gq.step_timing_drain(g)[0].size == 0
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_block_processing_mirror(ctx):
    """pipeline_process(RD) + process(UpdateTumor) == one fused step."""
    n = 32
    rng = np.random.default_rng(3)
    phi = (0.4 * rng.random((n, n, n))).astype(np.float32)
    pw = rng.random((n, n, n)).astype(np.float32)
    pg = (1 - pw).astype(np.float32)
    outs = []
    for mode in range(2):
        g = gq.Grid(ctx, 2, 1, 2)
        for name, v in (("phi", phi), ("pw", pw), ("pg", pg)):
            g.sample_volume(name, v)
        if mode == 0:
            bp = gq.BlockProcessing()
            vinfo = g.getBlocksInfo()
            bp.pipeline_process(vinfo, g, None, gq.Glioma_ReactionDiffusionOperator(1e-5, 1e-6, 0.03))
            bp.process(vinfo, g, gq.UpdateTumor(0.5))
        else:
            gq.fk_steps(g, 1e-5, 1e-6, 0.03, 0.5, 1)
        outs.append(g.dump_uniform())
    assert np.array_equal(outs[0], outs[1])


def test_step_variants_bit_identical():
    """Every step configuration the autotuner may pick (slab depths 16/8/4/2
    and the column-streaming kernel, each with the cross-level face ghosts
    from a separate launch or evaluated inside the step) gives the same bits
    on an adaptive grid with cross-level faces, so the tuning never changes
    results."""
    n = 64
    pwm, pgm, _ = uniform_case(n, seed=4)
    # a tumour inside one 32^3 octant: the other seven compress to level 1
    u0 = A.initial_condition(n, (n // 4, n // 4, n // 4), 0.6, 3.0)
    outs = []
    for mode, zt, pre in [(m, z, True) for m in (2, 3) for z in (16, 8, 4, 2, 1)] + [(2, 4, False), (2, 1, False)]:
        if True:
            ctx = gq.Context(0)
            gq.set_autotune(ctx, mode, zt)
            gq.set_step_coeffs(ctx, pre)
            g = gq.Grid(ctx, 4, 3, 4)
            for name, v in (("phi", u0), ("pw", pwm), ("pg", pgm)):
                g.write_volume(name, v)
            gq.Science.AutomaticCompression(g, 1e-3, -1)  # coarse leaves far from the tumour
            Dw = 0.0013 / L_CM ** 2 * 40
            h = 1.0 / n
            gq.fk_steps(g, Dw, 0.1 * Dw, 0.3, 0.9 * h * h / (6 * Dw), 7)
            assert gq.autotune_info(ctx)[:2] == (zt, mode == 3)
            outs.append((g.n_blocks(), g.read("phi")))
            ctx.close()
    assert 8 < outs[0][0] < 64  # the grid is adaptive
    for nb, a in outs[1:]:
        assert nb == outs[0][0] and np.array_equal(a, outs[0][1])


def test_persistent_chunks_bit_identical():
    """Chunks of steps run as one persistent cooperative launch (grid-wide
    barriers between the face-ghost and step phases) give the same bits as
    one launch per step, on an adaptive grid with cross-level faces, at slab
    depths 16, 4 and 2, with the face ghosts from the separate phase or
    evaluated in the step, for a chunk longer than one launch's 16 steps."""
    n = 64
    pwm, pgm, _ = uniform_case(n, seed=4)
    u0 = A.initial_condition(n, (n // 4, n // 4, n // 4), 0.6, 3.0)
    outs = []
    for persist, mode, zt in [(0, 2, 4), (2, 2, 16), (2, 2, 4), (2, 2, 2), (2, 3, 4), (1, 1, 16)]:
        ctx = gq.Context(0)
        gq.set_autotune(ctx, mode, zt)
        gq.set_step_persistent(ctx, persist)
        g = gq.Grid(ctx, 4, 3, 4)
        for name, v in (("phi", u0), ("pw", pwm), ("pg", pgm)):
            g.write_volume(name, v)
        gq.Science.AutomaticCompression(g, 1e-3, -1)
        Dw = 0.0013 / L_CM ** 2 * 40
        h = 1.0 / n
        gq.fk_steps(g, Dw, 0.1 * Dw, 0.3, 0.9 * h * h / (6 * Dw), 21)
        used, us = gq.persistent_info(ctx)
        print("persistent mode %d autotune %d zt %d: used %s, tuned %.1f us/step %s"
              % (persist, mode, zt, used, us, gq.autotune_info(ctx)[2]))
        if persist != 1:  # mode 1 keeps whichever the autotuner measured faster
            assert used == (persist == 2)
        if mode in (2, 3):  # the fixed slab depth holds whatever the persistent mode
            assert gq.autotune_info(ctx)[:2] == (zt, mode == 3)
        outs.append((g.n_blocks(), g.read("phi")))
        ctx.close()
    assert 8 < outs[0][0] < 64
    for nb, a in outs[1:]:
        assert nb == outs[0][0] and np.array_equal(a, outs[0][1])


def awkward_case(n, seed=11):
    """Anatomy and field with every case the per-cell 1/K codes treat apart."""
    rng = np.random.default_rng(seed)
    pw = rng.random((n, n, n)).astype(np.float32)
    pg = (1 - pw).astype(np.float32)
    phi = (0.5 * rng.random((n, n, n))).astype(np.float32)
    flat = [a.reshape(-1) for a in (pw, pg, phi)]
    idx = rng.permutation(n ** 3)
    k = 0
    for w_, g_, p_ in [(-0.01, 0.4, None), (0.3, -0.02, None), (0.2, -0.2, None), (-0.1, -0.05, None),
                       (0.0, 0.0, 0.3), (0.0, 0.0, 0.0), (0.0, 0.0, -0.01), (None, None, -0.02),
                       (-0.0, 0.0, 0.1), (1e-30, 0.0, None)]:
        sel = idx[k:k + 400]
        k += 400
        if w_ is not None:
            flat[0][sel] = w_
            flat[1][sel] = g_
        if p_ is not None:
            flat[2][sel] = p_
    return pw, pg, phi


@pytest.mark.parametrize("n,steps", [(32, 7), (64, 6)])
def test_fused_steps_bit_identical(n, steps):
    """Two steps fused into one launch (uniform grids) give the same bits as
    one launch per step, on anatomy with every special case of the 1/K codes
    and a field with zero regions, for odd and even step counts."""
    pw, pg, phi = awkward_case(n)
    phi[:, :, : n // 2] = 0.0  # a zero half: the closed-form zero update at its edge
    res = []
    for fuse in (False, True):
        c = gq.Context(0)
        gq.set_step_fusion(c, fuse)
        g = gq.Grid(c, n // 16, int(np.log2(n // 16)), 2)
        for name, v in (("phi", phi), ("pw", pw), ("pg", pg)):
            g.sample_volume(name, v)
        gq.fk_steps(g, 2e-4, 2e-5, 0.3, 0.5, steps)
        res.append(g.dump_uniform())
        c.close()
    assert np.isfinite(res[0]).all() and np.count_nonzero(res[0]) > 0
    assert res[0].tobytes() == res[1].tobytes()


@pytest.mark.parametrize("zt", [16, 4, 1])
def test_precomputed_coefficients_bit_identical(ctx, zt):
    """The per-cell 1/K codes give the same bits as recomputing 1/K from pw,
    pg every step, also for the cells that take the slow path: negative pw or
    pg (as Interp4 prolongation makes them), p_w = -p_g != 0, p_w + p_g < 0,
    CSF cells with and without tumour, and negative phi (as after a split)."""
    n = 32
    pw, pg, phi = awkward_case(n)
    res = []
    for pre in (False, True):
        c = gq.Context(0)
        gq.set_autotune(c, 2, zt)
        gq.set_step_coeffs(c, pre)
        g = gq.Grid(c, 2, 1, 2)
        for name, v in (("phi", phi), ("pw", pw), ("pg", pg)):
            g.sample_volume(name, v)
        gq.fk_steps(g, 2e-4, 2e-5, 0.3, 0.5, 3)
        res.append(g.dump_uniform())
        c.close()
    assert np.isfinite(res[0]).all()
    assert res[0].tobytes() == res[1].tobytes()


def test_neumann_doubling_matches_oracle(ctx):
    """Dw = 0 with white-matter-only cells (pw = 1, pg = 0: D = 0, 1/K = inf)
    next to grey-matter cells: the faces between them get df = 0, and the
    reference doubles the opposite face (Glioma_ReactionDiffusionOperator.h:
    198-203).  FP32 field against the FP64 oracle."""
    n = 32
    rng = np.random.default_rng(17)
    pw = (rng.random((n, n, n)) < 0.3).astype(np.float32)
    pg = (1 - pw).astype(np.float32)
    phi = (0.6 * rng.random((n, n, n))).astype(np.float32)
    Dw, Dg, rho, dt, steps = 0.0, 2e-5, 0.05, 0.9 / (n * n * 6 * 2e-5), 4
    g = gq.Grid(ctx, 2, 1, 2)
    for name, v in (("phi", phi), ("pw", pw), ("pg", pg)):
        g.sample_volume(name, v)
    gq.fk_steps(g, Dw, Dg, rho, dt, steps)
    u = g.dump_uniform()
    ref = O.fk_uniform_steps(phi.astype(np.float64), pw.astype(np.float64), pg.astype(np.float64), 16, 1.0 / n,
                             np.float32(Dw), np.float32(Dg), np.float32(rho), dt, steps, mode="f64")
    assert rel_l2(u, ref) <= 1e-5 and np.max(np.abs(u - ref)) <= 1e-5


@pytest.mark.parametrize("bpd", [1, 2])
def test_tiny_periodic_grids_and_zero_steps(bpd):
    """Degenerate sizes: one block (its own neighbour across every face of the
    periodic domain) and two blocks per dimension, at every slab depth, against
    the FP64 oracle; zero steps leave the field unchanged to the bit."""
    n = 16 * bpd
    rng = np.random.default_rng(bpd)
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    pw = (0.5 + 0.4 * np.sin(2 * np.pi * x / n) * np.cos(2 * np.pi * z / n)).astype(np.float32)
    pg = (1 - pw).astype(np.float32)
    pw[2:4, 5:7, 1:3] = 0  # a CSF pocket
    pg[2:4, 5:7, 1:3] = 0
    phi = (0.4 * np.exp(-((x - n / 2) ** 2 + (y - 3) ** 2 + (z - n / 2) ** 2) / 10.0)).astype(np.float32)
    phi += (0.01 * rng.random(phi.shape)).astype(np.float32)
    h = 1.0 / n
    Dw = 2e-4
    Dg, rho = 0.1 * Dw, 0.4
    dt = 0.9 * h * h / (6 * Dw)
    steps = 9
    ref = O.fk_uniform_steps(phi.astype(np.float64), pw.astype(np.float64), pg.astype(np.float64), 16, h, Dw, Dg,
                             rho, dt, steps)
    for zt in (16, 8, 4, 2, 1):
        ctx = gq.Context(0)
        gq.set_autotune(ctx, 2, zt)
        g = gq.Grid(ctx, bpd, int(np.log2(bpd)), 4)
        for name, v in (("phi", phi), ("pw", pw), ("pg", pg)):
            g.sample_volume(name, v)
        u_before = g.dump_uniform()
        gq.fk_steps(g, Dw, Dg, rho, dt, 0)
        assert g.dump_uniform().tobytes() == u_before.tobytes()
        gq.fk_steps(g, Dw, Dg, rho, dt, steps)
        u = g.dump_uniform()
        e2 = rel_l2(u, ref)
        print("bpd %d zt %d: rel L2 %.3e" % (bpd, zt, e2))
        assert e2 <= 1e-5 and float(np.max(np.abs(u - ref))) <= 1e-5
        assert zero_mismatch(u, ref) == 0
        ctx.close()
