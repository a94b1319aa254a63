"""Spiral FDK (fdk_spiral.cuh:28-94, spiral_ct_geom.h:30-95): per-slice view
windows on the host (CPU tests) and the reconstruction on the GPU against the
oracle's restatement (FP32 parity at the 1e-5 comparator bar, FP64 accuracy)."""
import math

import numpy as np
import pytest

import gct_amd as g
import oracle_lib as O
from gct_amd import workloads as W


@pytest.mark.parametrize("kind", [g.CYLINDRICAL, g.FLAT], ids=lambda k: g.KIND_NAMES[k])
def test_spiral_windows_match_oracle(kind):
    wl = W.spiral_case(kind)
    got = g.spiral_windows(wl.scan, wl.img)
    assert np.array_equal(got, O.spiral_windows(wl.scan, wl.img))
    assert (got[:, 1] > got[:, 0]).any()


def test_spiral_window_closed_form():
    """A slice at the source height of view m is centred on view m, and its
    window is the nppr views [m - nppr/2, m + nppr/2)."""
    wl = W.spiral_case(g.CYLINDRICAL, views_per_rot=64)
    r, _, _ = g.rotation_geom(wl.scan)
    nppr = r.num_proj_per_rot
    zsrc = wl.scan.views[:, 7]
    V = wl.scan.num_proj
    for m in (nppr // 2 + 1, V // 2, V - nppr // 2 - 1):
        img = g.Image(8, 8, 0.5, 0.5, 1, 0.5, mid=(0.0, 0.0, float(zsrc[m])))
        w = g.spiral_windows(wl.scan, img)[0]
        assert (w[0], w[1]) == (m - nppr // 2, m - nppr // 2 + nppr)
    # the reference skips a slice whose window would start at view 0 or run past the end
    for m in (nppr // 2, V - 1):
        img = g.Image(8, 8, 0.5, 0.5, 1, 0.5, mid=(0.0, 0.0, float(zsrc[m])))
        assert tuple(g.spiral_windows(wl.scan, img)[0]) == (0, 0)


def test_invalid_spirals_are_rejected():
    wl = W.spiral_case(g.CYLINDRICAL)
    # no pitch: a circular scan (spiral pitch cannot be zero)
    circ = g.circular_scan(g.CYLINDRICAL, 96, 0.5 / 200.0, 192, 2 * math.pi / 64, nv=16, dv=1.0,
                           src_det_dist=400.0, src_iso_dist=200.0)
    with pytest.raises(g.GctError, match="pitch"):
        g.spiral_windows(circ, wl.img)
    # source z not monotone
    bad = W.spiral_case(g.CYLINDRICAL).scan
    bad.views[10, 7], bad.views[11, 7] = bad.views[11, 7], bad.views[10, 7]
    with pytest.raises(g.GctError, match="sorted"):
        g.spiral_windows(bad, wl.img)
    # 1-D detectors have no spiral path
    fan = W.small_case(g.FAN_ARC_1D)
    with pytest.raises(g.GctError):
        g.spiral_windows(fan.scan, g.Image(8, 8, 0.5, 0.5, 4, 0.5))


def test_oracle_spiral_fp32_tracks_fp64():
    wl = W.spiral_case(g.CYLINDRICAL)
    proj = wl.projections()
    # a volume taller than the helix: the outer slices have no full rotation and stay zero
    img = g.Image(32, 32, 0.75, 0.75, 48, 0.5, mid=wl.img.mid)
    v32 = O.fdk_spiral(wl.scan, img, proj)
    v64 = O.fdk_spiral(wl.scan, img, proj, f64=True)
    win = O.spiral_windows(wl.scan, img)
    filled = win[:, 1] > win[:, 0]
    assert filled.sum() >= 8 and (~filled).sum() >= 8
    assert np.abs(v32[~filled]).max() == 0.0
    rep = O.parity_report(v32[filled], v64[filled])
    assert rep["max_rel_to_peak"] < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [g.CYLINDRICAL, g.FLAT], ids=lambda k: g.KIND_NAMES[k])
def test_spiral_reconstruction_matches_oracle(torch_dev, kind):
    wl = W.spiral_case(kind)
    proj = wl.projections()
    got = g.reconstruct_spiral(wl.scan, proj, wl.img)
    want = O.fdk_spiral(wl.scan, wl.img, proj)
    want64 = O.fdk_spiral(wl.scan, wl.img, proj, f64=True)
    rep = O.parity_report(got, want)
    err64 = float(np.sqrt(np.mean((got.astype(np.float64) - want64) ** 2)))
    print(f"\n[spiral {g.KIND_NAMES[kind]}] vs FP32 oracle {rep}; rmse vs FP64 {err64:.3e}", flush=True)
    assert O.are_equal(got, want, 1e-5).all()
    assert err64 < 1e-5
