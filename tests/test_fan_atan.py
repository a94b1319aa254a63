"""The FP32 fan-angle arctangent of the arc-detector kernels (bp_device.hpp:
fan_atan2) evaluated on the CPU in float32, step for step, against
numpy.arctan2 in FP64 over all eight octants and the axes.  The coefficients
are read from the kernel source, so a mistyped coefficient, a dropped term or a
wrong octant restore fails here before any GPU run."""
import pathlib
import re

import numpy as np

SRC = pathlib.Path(__file__).resolve().parents[1] / "gct_amd" / "csrc" / "bp_device.hpp"


def _coefs():
    m = re.search(r"kAtanCoef\[9\]\s*=\s*\{([^}]*)\}", SRC.read_text())
    assert m, "kAtanCoef not found"
    return np.array([float(s.strip().rstrip("f")) for s in m.group(1).split(",")], np.float32)


def _fan_atan2(y, x, c):
    y = y.astype(np.float32)
    x = x.astype(np.float32)
    ax, ay = np.abs(x), np.abs(y)
    mn, mx = np.minimum(ax, ay), np.maximum(ax, ay)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(mx > 0, mn * (np.float32(1) / mx), np.float32(0)).astype(np.float32)
    u = t * t
    p = np.full_like(t, c[8])
    for j in range(7, -1, -1):
        p = (p * u + c[j]).astype(np.float32)
    r = t * p
    r = np.where(ay > ax, np.float32(1.57079637) - r, r)
    r = np.where(x < 0, np.float32(3.14159274) - r, r)
    return np.copysign(r, y).astype(np.float32)


def test_fan_atan2_matches_arctan2():
    c = _coefs()
    rng = np.random.default_rng(7)
    ang = rng.uniform(-np.pi, np.pi, 400000)
    rad = rng.uniform(1e-3, 1e3, ang.size)
    y, x = (rad * np.sin(ang)).astype(np.float32), (rad * np.cos(ang)).astype(np.float32)
    got = _fan_atan2(y, x, c).astype(np.float64)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    err = np.abs(got - ref)
    # 1.2e-7 from the fit evaluated in FP32, plus the octant restore's
    # rounding of values up to pi (half an ulp of 3.14 is 1.2e-7).
    assert err.max() < 4e-7, err.max()


def test_fan_atan2_axes_and_origin():
    c = _coefs()
    y = np.array([0, 0, 1, -1, 0, 0, 2, -2], np.float32)
    x = np.array([0, 1, 0, 0, -1, 3, 2, 2], np.float32)
    got = _fan_atan2(y, x, c)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    np.testing.assert_allclose(got, ref, atol=3e-7)
    assert got[0] == 0.0


def test_fit_script_reproduces_header():
    import importlib.util

    spec = importlib.util.spec_from_file_location("atan_fit", SRC.parent / "atan_fit.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    c = mod.fit().astype(np.float32)
    np.testing.assert_allclose(c, _coefs(), rtol=2.4e-7)   # within an ulp of the header


def _narrow():
    txt = SRC.read_text()
    m = re.search(r"kAtanNarrowCoef\[7\]\s*=\s*\{([^}]*)\}", txt)
    t = re.search(r"kAtanNarrowT\s*=\s*([0-9.eE+-]+)f", txt)
    assert m and t, "narrow arctangent constants not found"
    return (np.array([float(s.strip().rstrip("f")) for s in m.group(1).split(",")], np.float32),
            np.float32(float(t.group(1))))


def _fan_atan_narrow(nu, nw, c, T):
    """bp_device.hpp fan_atan_narrow, step for step in float32 (the reciprocal
    is IEEE here, within an ulp of the kernel's v_rcp_f32)."""
    nu = nu.astype(np.float32)
    nw = nw.astype(np.float32)
    t = np.clip((nu * (np.float32(1) / nw)).astype(np.float32), -T, T)
    u = (t * t).astype(np.float32)
    p = np.full_like(t, c[6])
    for j in range(5, -1, -1):
        p = (p * u + c[j]).astype(np.float32)
    return (t * p).astype(np.float32)


def test_fan_atan_narrow_matches_arctan_in_range():
    c, T = _narrow()
    assert abs(float(T) - np.tan(0.6)) < 1e-7
    rng = np.random.default_rng(11)
    gam = rng.uniform(-0.6, 0.6, 400000)
    rad = rng.uniform(1.0, 1e3, gam.size)
    nu, nw = (rad * np.sin(gam)).astype(np.float32), (rad * np.cos(gam)).astype(np.float32)
    got = _fan_atan_narrow(nu, nw, c, T).astype(np.float64)
    ref = np.arctan2(nu.astype(np.float64), nw.astype(np.float64))
    # 7.9e-8 from the fit evaluated in FP32, plus the rounding of t = nu / nw
    # (half an ulp of the quotient and of the product) carried through atan'
    assert np.abs(got - ref).max() < 1.6e-7


def test_fan_atan_narrow_clamps_beyond_its_range():
    """Angles beyond 0.6 rad take the clamped angle (0.6 within the fit), with
    the right sign: a voxel outside the fan still lands outside the detector."""
    c, T = _narrow()
    gam = np.array([0.61, 0.9, 1.4, -0.61, -1.2], np.float64)
    nu, nw = np.sin(gam).astype(np.float32), np.cos(gam).astype(np.float32)
    got = _fan_atan_narrow(nu, nw, c, T)
    np.testing.assert_allclose(np.abs(got), 0.6, atol=1.2e-7)
    assert np.all(np.sign(got) == np.sign(gam))


def test_fit_script_reproduces_narrow_header():
    import importlib.util

    spec = importlib.util.spec_from_file_location("atan_fit", SRC.parent / "atan_fit.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    c, T = _narrow()
    np.testing.assert_allclose(mod.fit_narrow().astype(np.float32), c, rtol=2.4e-7)
    assert np.float32(mod.NARROW_T) == T
