"""Voxel-to-detector mapping and backprojection weights, pinned per voxel.

A single view's detector coordinates and weight are read back from three
unscaled backprojections of filtered stacks that are zero except in that
view: all ones (gives the weight w), the column index (w * column) and the
row index (w * row).  Bilinear interpolation reproduces a linear ramp
exactly, so column and row come out in detector pixels.  The same probe runs
on the oracle (CPU) and on the GPU library.

Pinned against:
* the reference's own known answers for the mapping
  (test/cuda/geom/coord_trans_proj_test.cu:9-395: the fixtures of
  test/cpp/geom/proj_geom_params_test_fix.h -- 4 views from pi/2 in steps of
  pi/2, view 2, a 512 x 512 detector of unit pitch centred at (256, 257),
  SID 500, SDD 1000 -- world origin to the detector centre, weights 1, 0.002,
  2.0, sign symmetry) and for voxel centres
  (test/cuda/geom/coord_trans_img_test.cu:6-68);
* an FP64 evaluation, written here from the published projection formulas
  (coord_trans_proj.cuh:35-238), of column, row and weight for off-centre
  voxels in one view of a generic geometry -- one view, so that a wrong
  distance exponent cannot average out over a rotation; the test also shows
  that the plausible wrong forms (another exponent, |r| for r.ew, r.ew for
  |r|) miss by far more than the tolerance.
"""
import math

import numpy as np
import pytest

import gct_amd as g
import oracle_lib as O


def _oracle_bp(scan, img, filt):
    return O.backproject(scan, img, filt, scale=False, nthreads=4).astype(np.float64)


def _gpu_bp(scan, img, filt):
    import torch
    plan = g.Plan(scan, img)
    plan.load_filtered(torch.from_numpy(np.ascontiguousarray(filt)).cuda())
    v = torch.zeros(img.shape, dtype=torch.float32, device="cuda")
    plan.backproject(v, finish=False)
    torch.cuda.synchronize()
    plan.close()
    return v.cpu().numpy().astype(np.float64)


BACKENDS = [pytest.param(_oracle_bp, id="oracle"), pytest.param(_gpu_bp, id="gpu", marks=pytest.mark.gpu)]


def probe(bp, scan, img, view):
    """(column, row, weight) of every voxel of img in view `view`."""
    shape = scan.proj_shape

    def stack(plane):
        f = np.zeros(shape, np.float32)
        f[view] = plane
        return f

    ones = stack(np.ones(shape[1:], np.float32))
    cols = stack(np.broadcast_to(np.arange(scan.nu, dtype=np.float32), shape[1:]))
    rows = stack(np.broadcast_to(np.arange(scan.nv, dtype=np.float32)[:, None], shape[1:]))
    w = bp(scan, img, ones)
    return bp(scan, img, cols) / w, bp(scan, img, rows) / w, w


# ---- the reference's known answers ------------------------------------------------

def kat_scan(kind):
    """The fixtures of proj_geom_params_test_fix.h: CW, init pi/2, increment
    pi/2, 4 views; detector 512 (x 512) of unit pitch, centre (256, 257);
    SID 500 mm, SDD 1000 mm."""
    pi = np.float32(math.pi)
    return g.circular_scan(kind, 512, 1.0, 4, float(pi / np.float32(2.0)), init_ang=float(pi / np.float32(2.0)),
                           nv=512, dv=1.0, det_row_center=256.0, det_col_center=257.0, src_det_dist=1000.0,
                           src_iso_dist=500.0)


def point(kind, p):
    if g.is_1d(kind):
        return g.Image(1, 1, 1.0, 1.0, mid=tuple(p[:2]))
    return g.Image(1, 1, 1.0, 1.0, 1, 1.0, mid=tuple(p))


KAT_WEIGHT = {g.PARALLEL_1D: 1.0, g.FAN_ARC_1D: 0.002, g.FAN_LINE_1D: 2.0, g.PARALLEL_2D: 1.0,
              g.CONE_PARALLEL: 1.0, g.CYLINDRICAL: 0.002, g.FLAT: 2.0}


@pytest.mark.parametrize("bp", BACKENDS)
@pytest.mark.parametrize("kind", range(7), ids=g.KIND_NAMES)
def test_reference_mapping_known_answers(bp, kind):
    scan = kat_scan(kind)
    _, rot, _ = g.rotation_geom(scan)
    assert float(rot[2]) == pytest.approx(4.7123889923, rel=1e-6)   # view 2 at 3 pi / 2
    col, row, w = (float(a.ravel()[0]) for a in probe(bp, scan, point(kind, (0.0, 0.0, 0.0)), 2))
    # world origin -> the detector centre; the reference's weights at the origin
    assert col == pytest.approx(256.0, abs=1e-4)
    if not g.is_1d(kind):
        assert row == pytest.approx(257.0, abs=1e-4)
    assert w == pytest.approx(KAT_WEIGHT[kind], rel=1e-5)
    pos = [float(a.ravel()[0]) for a in probe(bp, scan, point(kind, (2.0, 0.0, 2.0)), 2)]
    neg = [float(a.ravel()[0]) for a in probe(bp, scan, point(kind, (-2.0, 0.0, -2.0)), 2)]
    assert pos[0] > neg[0]
    assert pos[0] - 256.0 == pytest.approx(256.0 - neg[0], rel=1e-4)
    if kind == g.PARALLEL_1D:
        # detCoordAxialParallel at 3 pi / 2: (+-2, 0) -> +-2 mm
        assert pos[0] == pytest.approx(258.0, abs=1e-4) and neg[0] == pytest.approx(254.0, abs=1e-4)
    if not g.is_1d(kind):
        assert pos[1] < neg[1]
        assert 257.0 - pos[1] == pytest.approx(neg[1] - 257.0, rel=1e-4)


@pytest.mark.parametrize("bp", BACKENDS)
def test_reference_voxel_centres(bp):
    """Voxel centre x_i = mid - (N - 1) d / 2 + i d (coord_trans_img_test.cu:
    1025 x 513 x 11 voxels of 1 x 2 x 4 mm at (1, 2, 3): first voxel
    (-511, -510, -17)), read back through 2-D parallel views at 0 and pi/2,
    where the column is y (then -x) over the pitch plus the centre and the
    row is -z over the pitch plus the centre."""
    scan = g.circular_scan(g.PARALLEL_2D, 1200, 1.0, 2, math.pi / 2, nv=64, dv=1.0, det_row_center=600.0,
                           det_col_center=32.0)
    img = g.Image(1025, 513, 1.0, 2.0, 11, 4.0, mid=(1.0, 2.0, 3.0))
    c0, r0, _ = probe(bp, scan, img, 0)
    c1, _, _ = probe(bp, scan, img, 1)
    x = 600.0 - c1[0, 0, :]
    y = c0[0, :, 0] - 600.0
    z = 32.0 - r0[:, 0, 0]
    assert (x[0], y[0], z[0]) == pytest.approx((-511.0, -510.0, -17.0), abs=2e-3)
    np.testing.assert_allclose(x, 1.0 - 1024 * 1.0 / 2 + np.arange(1025) * 1.0, atol=2e-3)
    np.testing.assert_allclose(y, 2.0 - 512 * 2.0 / 2 + np.arange(513) * 2.0, atol=2e-3)
    np.testing.assert_allclose(z, 3.0 - 10 * 4.0 / 2 + np.arange(11) * 4.0, atol=2e-3)


# ---- FP64 evaluation of the projection formulas -------------------------------------

def generic_scan(kind):
    """Off-centre detector, unequal pitches, a view at a generic angle."""
    pitch_u = 0.0021 if kind in (g.FAN_ARC_1D, g.CYLINDRICAL) else 0.7
    return g.circular_scan(kind, 700, pitch_u, 5, 0.9, init_ang=0.31, nv=500, dv=0.55, det_row_center=341.3,
                           det_col_center=262.6, src_det_dist=1100.0, src_iso_dist=600.0)


def expected(scan, img, view):
    """(column, row, weight) per voxel in FP64 from the view's record:
    1-D parallel t = x cos(th + pi/2) + y sin(th + pi/2), w = 1; 2-D parallel
    adds row -z; cone-parallel row -z D / SID; fan arc gamma =
    atan2(r.eu, r.ew), w = SID / |r|^2; fan line u = D (r.eu) / (r.ew),
    w = D SID / (r.ew)^2 with ew = rot90(eu); cylindrical gamma as the arc,
    v = D (r^.ev) / sqrt(1 - (r^.ev)^2), w = SID / |r|^2; flat u, v =
    D (r.eu, r.ev) / (r.ew), ew = eu x ev, w = D SID / (r.ew)^2.  SID = |src|
    except for cone-parallel (its parameter).  Column = u / du + cu, row =
    v / dv + cv.  Also returns the weights plausible mistakes would give."""
    rec = scan.views[view].astype(np.float64)
    th, D, sid_par, cu, cv = rec[0], rec[1], rec[2], rec[3], rec[4]
    src, eu, ev, ew_f = rec[5:8], rec[8:11], rec[11:14], rec[14:17]
    nz = img.nz if img.dim == 3 else 1
    ax = [img.mid[i] - ((n - 1) * d) / 2 + np.arange(n) * d
          for i, (n, d) in enumerate(((img.nx, img.dx), (img.ny, img.dy), (nz, img.dz)))]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    if img.dim == 2:
        Z = np.zeros_like(Z)
    kind = scan.kind
    wrong = []
    row = np.zeros_like(X)
    if kind in (g.PARALLEL_1D, g.PARALLEL_2D, g.CONE_PARALLEL):
        t = X * math.cos(th + math.pi / 2) + Y * math.sin(th + math.pi / 2)
        col = t / scan.du + cu
        if kind == g.PARALLEL_2D:
            row = -Z / scan.dv + cv
        elif kind == g.CONE_PARALLEL:
            row = -Z * D / sid_par / scan.dv + cv
            wrong = [np.full_like(X, D / sid_par)]
        w = np.ones_like(X)
        return col, row, w, wrong
    if g.is_1d(kind):
        src, eu, ev, ew_f = src * [1, 1, 0], eu * [1, 1, 0], ev * 0, ew_f * [1, 1, 0]
        Z = np.zeros_like(Z)
    sid = float(np.linalg.norm(src))
    r = np.stack([X - src[0], Y - src[1], Z - src[2]], -1)
    if kind == g.FAN_LINE_1D:
        ew = np.array([-eu[1], eu[0], 0.0])
    elif kind == g.FLAT:
        ew = np.cross(eu, ev)
    else:
        ew = ew_f
    ru, rv, rw = r @ eu, r @ ev, r @ ew
    rn2 = (r * r).sum(-1)
    if kind in (g.FAN_ARC_1D, g.CYLINDRICAL):
        col = np.arctan2(ru, rw) / scan.du + cu
        if kind == g.CYLINDRICAL:
            c = rv / np.sqrt(rn2)
            row = D * c / np.sqrt(1 - c * c) / scan.dv + cv
        w = sid / rn2
        wrong = [sid / np.sqrt(rn2), sid / rn2 ** 1.5, sid / rw ** 2, 1.0 / rn2]
    else:
        col = D * ru / rw / scan.du + cu
        if kind == g.FLAT:
            row = D * rv / rw / scan.dv + cv
        w = D * sid / rw ** 2
        wrong = [D * sid / np.abs(rw), D * sid / np.abs(rw) ** 3, D * sid / rn2, D / rw ** 2, sid / rw ** 2]
    return col, row, w, wrong


@pytest.mark.parametrize("bp", BACKENDS)
@pytest.mark.parametrize("kind", range(7), ids=g.KIND_NAMES)
def test_mapping_and_weight_match_fp64_formulas(bp, kind):
    scan = generic_scan(kind)
    if g.is_1d(kind):
        img = g.Image(5, 4, 23.0, 31.0, mid=(41.0, -57.0))
    else:
        img = g.Image(5, 4, 23.0, 31.0, 3, 37.0, mid=(41.0, -57.0, 19.0))
    view = 3
    col, row, w = probe(bp, scan, img, view)
    ecol, erow, ew, wrong = expected(scan, img, view)
    assert np.all((ecol > 2) & (ecol < scan.nu - 3))
    np.testing.assert_allclose(w, ew.reshape(w.shape), rtol=2e-5)
    np.testing.assert_allclose(col, ecol.reshape(col.shape), atol=2e-3)
    if not g.is_1d(kind):
        assert np.all((erow > 2) & (erow < scan.nv - 3))
        np.testing.assert_allclose(row, erow.reshape(row.shape), atol=2e-3)
    # every plausible wrong weight misses somewhere by far more than the tolerance
    for alt in wrong:
        rel = np.abs(alt.reshape(w.shape) / w - 1.0)
        assert rel.max() > 100 * 2e-5, "a wrong weight form would pass this test"
