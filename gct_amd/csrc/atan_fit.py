"""Coefficients of the fan-angle arctangents used by the FP32 kernels
(bp_device.hpp): atan(t) ~ t * P(t^2), fitted by least squares on 4000
Chebyshev nodes in FP64.  Prints the coefficients (lowest degree first) and
each fit's largest error on a dense grid.

  fan_atan2:       t in [0, 1] (after octant reduction), P of degree 8:
                   6.3e-9 in FP64; evaluated in FP32 by Horner's rule, 1.2e-7
                   (about 1.5 ulp at pi/4).
  fan_atan_narrow: t in [0, tan 0.6] (fans within +-0.6 rad, every voxel in
                   front of the source: no octant reduction), P of degree 6:
                   3.3e-9 in FP64, 7.9e-8 in FP32.

    python gct_amd/csrc/atan_fit.py
"""
import math

import numpy as np

DEG = 8
NARROW_DEG = 6
NARROW_T = float(np.float32(math.tan(0.6)))   # the kernel's clamp, an FP32 value


def fit(deg=DEG, nodes=4000, T=1.0):
    k = np.arange(nodes)
    t = T * 0.5 * (1.0 - np.cos(np.pi * (k + 0.5) / nodes))
    a = np.stack([t ** (2 * j + 1) for j in range(deg + 1)], 1)
    c, *_ = np.linalg.lstsq(a, np.arctan(t), rcond=None)
    return c


def fit_narrow(nodes=4000):
    return fit(NARROW_DEG, nodes, NARROW_T)


def max_errors(c, T=1.0):
    tt = np.linspace(0.0, T, 200001)
    exact = np.arctan(tt)
    f64 = sum(cj * tt ** (2 * j + 1) for j, cj in enumerate(c))
    c32 = c.astype(np.float32)
    t32 = tt.astype(np.float32)
    u = t32 * t32
    p = np.float32(c32[-1])
    for cj in c32[-2::-1]:
        p = np.float32(p * u + cj)
    return np.max(np.abs(f64 - exact)), np.max(np.abs((t32 * p).astype(np.float64) - exact))


if __name__ == "__main__":
    for name, c, T in (("kAtanCoef", fit(), 1.0), ("kAtanNarrowCoef", fit_narrow(), NARROW_T)):
        e64, e32 = max_errors(c, T)
        print("constexpr float %s[%d] = {%s};" % (name, len(c), ", ".join("%.9ef" % float(np.float32(x)) for x in c)))
        print("max error on [0, %.9f]: FP64 %.2e, FP32 Horner %.2e" % (T, e64, e32))
