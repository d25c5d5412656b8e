"""Mathematical pins of the oracle functions whose reference code cannot be
compiled here (MRAGBlockFWT.h, MRAGBlockSplitter.h and Glioma_HG_UQ.cpp all
include MRAGrid.h, which needs the TBB headers):

* fwt_block_detail_f64 (BlockFWT::_filter/_fwt + FWTReport, MRAGBlockFWT.h:
  49-192, 220-298): interpolating-wavelet details vanish on polynomials of
  degree < order in each variable; on x^order and x^(order+1) they equal the
  Lagrange interpolation remainder f[x_0..x_k, x] * prod(x - x_j), which fixes
  the node positions and the sign of the stencil; single spikes give the Ga
  taps 1, -9/16, 0, 1/16 (Interp4, MRAGWavelets_StaticData.cpp:32) and 1, -1/2
  (Interp2, :14) -- the tap at the odd point, the ones next to it, the zero
  tap and the outermost tap; the scaling coefficients (code 0) are excluded
  from the block's maximum (FWTReport::_getDetailMaxMag, MRAGBlockFWT.h:
  109-120).
* AdaptiveField._split (BlockSplitter::_splitBlocks, MRAGBlockSplitter.h:24-99):
  the children of a split block reproduce cubics (Interp4) / linears
  (Interp2) sampled at the parent's points exactly, in the right child.
* dump_uniform_f64 (Glioma_HG_UQ::_dumpUQoutput, Glioma_HG_UQ.cpp:309-372):
  replication factors 1, 2, 4 and the holes a block three levels coarser
  leaves, from an independent covering-cell formulation.
* patient_ic (Glioma_HG_UQ::_ic_PatientCase, Glioma_HG_UQ.cpp:67-225): hand-
  derived special cases -- normalisation, CSF kept below 0.1 and made pure
  above it, CSF removed under MRI signal, the aspect-ratio shift, and the
  smoothed tumour ball (its closed-form value at the centre, its support and
  its absence outside tissue)."""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle import patient_oracle as P
from oracle import topology as T

B = 16


def lab_of(f, order):
    """Lab (stencil of the order) holding f(x, y, z) at block coordinates."""
    s0, s1 = (-2, 4) if order == 4 else (-1, 2)
    r = np.arange(s0, B + s1 - 1, dtype=np.float64)
    z, y, x = np.meshgrid(r, r, r, indexing="ij")
    return f(x, y, z), s0, len(r)


def detail(f, order):
    lab, s0, L = lab_of(f, order)
    return O.lib().fwt_block_detail_f64(np.ascontiguousarray(lab), B, s0, L, order)


@pytest.mark.parametrize("order", [2, 4])
def test_fwt_annihilates_polynomials(order):
    rng = np.random.default_rng(order)
    deg = order - 1
    for _ in range(5):
        c = rng.normal(size=(deg + 1, deg + 1, deg + 1))

        def f(x, y, z):
            v = np.zeros_like(x)
            for i, j, k in itertools.product(range(deg + 1), repeat=3):
                v += c[i, j, k] * (x / 8) ** i * (y / 8) ** j * (z / 8) ** k
            return v

        scale = np.max(np.abs(lab_of(f, order)[0]))
        assert detail(f, order) <= 1e-13 * scale
    # one degree more in one variable is not annihilated
    assert detail(lambda x, y, z: (x / 8) ** order, order) > 1e-6


@pytest.mark.parametrize("order,k,expected", [
    # Lagrange remainder at the odd point 2d+1 with nodes 2d-2, 2d, 2d+2, 2d+4
    # (Interp4) or 2d, 2d+2 (Interp2): x^order -> prod(x - x_j); x^(order+1) ->
    # (sum of nodes + x) * prod(x - x_j), largest at the last coarse point d = 7
    (4, 4, 9.0), (4, 5, 9.0 * (10 * 7 + 5)),
    (2, 2, 1.0), (2, 3, 6 * 7 + 3.0),
])
def test_fwt_interpolation_remainder(order, k, expected):
    d = detail(lambda x, y, z: x ** k, order)
    assert d == pytest.approx(expected, rel=1e-12)
    # the same along y and z (codes 2 and 4)
    assert detail(lambda x, y, z: y ** k, order) == pytest.approx(expected, rel=1e-12)
    assert detail(lambda x, y, z: z ** k, order) == pytest.approx(expected, rel=1e-12)


def spike_detail(p, order):
    def f(x, y, z):
        return ((x == p[0]) & (y == p[1]) & (z == p[2])).astype(np.float64)

    return detail(f, order)


@pytest.mark.parametrize("order,point,expected", [
    (4, (5, 5, 5), 1.0),          # odd point: the unit tap of code 7
    (4, (4, 6, 8), 9 / 16.),      # even point: the -9/16 taps (code 0 = 1 is excluded)
    (4, (-2, 2, 2), 1 / 16.),     # reached only by the outermost tap of d = 0
    (4, (-1, 2, 2), 0.0),         # reached only by the zero tap
    (4, (18, 2, 2), 1 / 16.),     # outermost tap of d = 7 on the high side
    (4, (17, 2, 2), 0.0),         # zero tap of d = 7
    (2, (5, 5, 5), 1.0),
    (2, (4, 6, 8), 0.5),
    (2, (16, 2, 2), 0.5),         # last tap of d = 7
    (2, (-1, 2, 2), 0.0),         # outside every filter's reach
])
def test_fwt_single_spike(order, point, expected):
    assert spike_detail(point, order) == pytest.approx(expected, rel=1e-15, abs=0)


@pytest.mark.parametrize("order", [2, 4])
def test_fwt_excludes_scaling_coefficients(order):
    # a constant field: code 0 carries the constant, every detail is zero
    assert detail(lambda x, y, z: np.full_like(x, 3.5), order) == 0.0
    # the maximum runs over every code 1..7: a pure xyz-wavelet pattern
    # ((-1)^(x+y+z) on odd points only) shows up through code 7 alone
    v = detail(lambda x, y, z: ((x % 2) * (y % 2) * (z % 2)).astype(np.float64), order)
    assert v == pytest.approx(1.0, rel=1e-15)


def poly_field(order, rng):
    deg = order - 1
    c = rng.normal(size=(deg + 1, deg + 1, deg + 1))

    def f(x, y, z):
        v = np.zeros_like(x)
        for i, j, k in itertools.product(range(deg + 1), repeat=3):
            v += c[i, j, k] * (4 * x) ** i * (4 * y) ** j * (4 * z) ** k
        return v

    return f


@pytest.mark.parametrize("order", [2, 4])
def test_split_reproduces_polynomials(order):
    rng = np.random.default_rng(10 + order)
    topo = T.Topology(4, 3, B=B, order=order)
    fld = T.AdaptiveField(topo)
    fs = {name: poly_field(order, rng) for name in fld.FIELDS}

    def sampler(f):
        def vol(l, org, h):
            r = np.arange(B) * h
            z, y, x = np.meshgrid(org[2] + r, org[1] + r, org[0] + r, indexing="ij")
            return f(x, y, z)
        return vol

    for name in fld.FIELDS:
        fld.set_from_volume(name, sampler(fs[name]))
    parent = (2, 1, 2, 1)  # interior: its lab does not wrap the periodic boundary
    assert fld.refine({topo.leaves[parent]}) == 1
    kids = [k for k in topo.leaves if k[0] == 3]
    assert len(kids) == 8
    for k in kids:
        assert k[1] // 2 == parent[1] and k[2] // 2 == parent[2] and k[3] // 2 == parent[3]
        h = 2.0 ** -3 / B
        org = np.array(k[1:]) * 2.0 ** -3
        for name in fld.FIELDS:
            want = sampler(fs[name])(3, org, h)
            got = fld.data[topo.leaves[k]][name]
            assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), (k, name)
    # one degree more is not reproduced at the odd child points
    topo2 = T.Topology(4, 3, B=B, order=order)
    fld2 = T.AdaptiveField(topo2)
    for name in fld2.FIELDS:
        fld2.set_from_volume(name, sampler(lambda x, y, z: (4 * x) ** order))
    fld2.refine({topo2.leaves[parent]})
    k = (3, 2, 4, 2)
    want = sampler(lambda x, y, z: (4 * x) ** order)(3, np.array(k[1:]) * 2.0 ** -3, 2.0 ** -3 / B)
    assert np.max(np.abs(fld2.data[topo2.leaves[k]]["phi"] - want)) > 1e-6


def test_dump_replication_and_holes():
    """Leaves on levels 1..4 with B = 8 dumped to gpd = 128 (the level-4
    spacing): factors 1 (level 4), 2 (level 3), 4 (level 2, covers its cells)
    and 4 of 8 (level 1, leaves holes)."""
    Bs = 8
    keys = [(1,) + c for c in itertools.product((0, 1), repeat=3)]
    for l in (1, 2, 3):
        p = (l, 0, 0, 0)
        keys.remove(p)
        keys += [(l + 1, c[0], c[1], c[2]) for c in itertools.product((0, 1), repeat=3)]
    keys.sort()
    lv = np.array([k[0] for k in keys], np.int32)
    ix = np.array([k[1:] for k in keys], np.int32)
    grid = O.LeafGrid(lv, ix, Bs, 4, 4)
    rng = np.random.default_rng(3)
    phi = rng.uniform(0.5, 1.5, size=(len(keys), Bs, Bs, Bs))
    gpd = 128
    got = grid.dump_uniform(phi.reshape(len(keys), -1), gpd)
    want = np.zeros((gpd, gpd, gpd))
    written = np.zeros((gpd, gpd, gpd), bool)
    for b, k in enumerate(keys):
        s = 1 << (4 - k[0])          # finest voxels per cell of this leaf
        w = min(s, 4)                # voxels the reference writes per cell and axis
        base = np.array(k[1:]) * Bs * s
        for (cz, cy, cx) in itertools.product(range(Bs), repeat=3):
            z0, y0, x0 = base[2] + cz * s, base[1] + cy * s, base[0] + cx * s
            want[z0:z0 + w, y0:y0 + w, x0:x0 + w] = phi[b, cz, cy, cx]
            written[z0:z0 + w, y0:y0 + w, x0:x0 + w] = True
    np.testing.assert_array_equal(got, want)
    # level-1 cells cover 8^3 voxels of which 4^3 are written: 7/8 holes there
    lvl1 = np.zeros_like(written)
    lvl1[64:, :, :] = True
    assert written[lvl1].mean() == pytest.approx(1 / 8)
    assert written[:64, :64, :64].all() and (got[:64, :64, :64] > 0).all()


# ---- patient anatomy -------------------------------------------------------

def anatomy(shape, gm=0.0, wm=0.0, csf=0.0, t1=0.0, t2=0.0):
    return [np.full(shape, v, np.float64) for v in (gm, wm, csf, t1, t2)]


FAR = (0.9, 0.9, 0.9)  # tumour centre away from the probed points


def test_patient_normalisation_and_csf_rules():
    gpd = 16
    # grey 0.3, white 0.6, CSF 0.05 (normalised CSF 0.05/0.95 < 0.1: kept)
    phi, pw, pg = P.patient_ic(*anatomy((8, 8, 8), 0.3, 0.6, 0.05), FAR, gpd)
    all_ = 0.3 + 0.6 + 0.05
    assert pw[3, 4, 5] == pytest.approx(0.6 / all_, rel=1e-6)
    assert pg[3, 4, 5] == pytest.approx(0.3 / all_, rel=1e-6)
    # CSF 0.2 of 1.1: normalised 0.18 > 0.1 -> pure CSF, no tissue
    phi, pw, pg = P.patient_ic(*anatomy((8, 8, 8), 0.3, 0.6, 0.2), FAR, gpd)
    assert np.all(pw == 0) and np.all(pg == 0)
    # the same with MRI signal: CSF removed before normalising
    phi, pw, pg = P.patient_ic(*anatomy((8, 8, 8), 0.3, 0.6, 0.2, t1=1.0), FAR, gpd)
    assert pw[1, 2, 3] == pytest.approx(0.6 / 0.9, rel=1e-6)
    assert pg[1, 2, 3] == pytest.approx(0.3 / 0.9, rel=1e-6)
    # no tissue at all: zero anatomy, no tumour even at the tumour centre
    phi, pw, pg = P.patient_ic(*anatomy((8, 8, 8)), (0.5, 0.5, 0.5), gpd)
    assert not pw.any() and not pg.any() and not phi.any()


def test_patient_aspect_shift():
    """Anatomy of 4 x 8 x 8 voxels on the unit cube: voxel size 1/8 along the
    longest axis, x shifted by (8 - 4) / 2 = 2 voxels.  A marker at anatomy
    x-index 1 lands on grid points with floor(8 x) - 2 == 1, i.e. x in
    [3/8, 4/8); x-indices outside [0, 4) hold no tissue."""
    gpd = 32
    GM, WM, CSF, T1, T2 = anatomy((4, 8, 8), gm=1.0)
    WM[1, :, :] = 1.0
    GM[1, :, :] = 0.0
    phi, pw, pg = P.patient_ic(GM, WM, CSF, T1, T2, FAR, gpd)
    x = np.arange(gpd) / gpd
    mx = np.floor(8 * x).astype(int) - 2
    for i in range(gpd):
        col_pw, col_pg = pw[5, 7, i], pg[5, 7, i]
        if mx[i] < 0 or mx[i] >= 4:
            assert col_pw == 0 and col_pg == 0, i
        elif mx[i] == 1:
            assert col_pw == 1 and col_pg == 0, i
        else:
            assert col_pw == 0 and col_pg == 1, i


def test_patient_tumour_ball():
    """Radius 0.005 smoothed over two points of h = 1/128: psi = (r - 0.005) *
    64; phi = (1 - psi - sin(pi psi) / pi) / 2 on |psi| <= 1 in tissue, 0
    beyond; psi < -1 would need r < 0 (never)."""
    gpd = 128
    centre = (64 / 128, 40 / 128, 30 / 128)  # on a grid point
    phi, pw, pg = P.patient_ic(*anatomy((16, 16, 16), gm=1.0), centre, gpd)
    psi0 = (0.0 - 0.005) * 64
    want0 = 0.5 * (1 - psi0 - math.sin(math.pi * psi0) / math.pi)
    assert phi[30, 40, 64] == pytest.approx(want0, rel=1e-6)
    # one point away: r = 1/128
    psi1 = (1 / 128 - 0.005) * 64
    want1 = 0.5 * (1 - psi1 - math.sin(math.pi * psi1) / math.pi)
    assert phi[30, 40, 65] == pytest.approx(want1, rel=1e-5)
    assert phi[31, 40, 64] == pytest.approx(want1, rel=1e-5)
    # support: r <= 0.005 + 1/64, i.e. at most 2 points along an axis
    assert phi[30, 40, 67] == 0.0 and phi[30, 40, 61] == 0.0
    nz = np.argwhere(phi > 0)
    r = np.sqrt(((nz[:, ::-1] / 128.0 - np.array(centre)) ** 2).sum(1))
    assert r.max() <= 0.005 + 2 / 128 + 1e-7
    # outside tissue the ball is absent
    GM, WM, CSF, T1, T2 = anatomy((16, 16, 16), gm=1.0)
    GM[8, :, :] = 0.0  # anatomy x-index 8 covers grid x in [64, 72)
    phi2, _, _ = P.patient_ic(GM, WM, CSF, T1, T2, centre, gpd)
    assert phi2[30, 40, 64] == 0.0 and phi2[30, 40, 63] > 0
