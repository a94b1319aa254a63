"""Coefficients of the fan-angle arctangent used by the FP32 kernels
(bp_device.hpp: fan_atan2): atan(t) ~ t * P(t^2) on [0, 1], P of degree 8,
fitted by least squares on 4000 Chebyshev nodes in FP64.  Prints the
coefficients (lowest degree first) and the fit's largest error on a dense grid:
6.3e-9 in FP64; evaluated in FP32 by Horner's rule, 1.2e-7 (about 1.5 ulp at
pi/4).

    python gct_amd/csrc/atan_fit.py
"""
import numpy as np

DEG = 8


def fit(deg=DEG, nodes=4000):
    k = np.arange(nodes)
    t = 0.5 * (1.0 - np.cos(np.pi * (k + 0.5) / nodes))
    a = np.stack([t ** (2 * j + 1) for j in range(deg + 1)], 1)
    c, *_ = np.linalg.lstsq(a, np.arctan(t), rcond=None)
    return c


def max_errors(c):
    tt = np.linspace(0.0, 1.0, 200001)
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
    c = fit()
    e64, e32 = max_errors(c)
    print("constexpr float kAtanCoef[%d] = {%s};" % (len(c), ", ".join("%.9ef" % float(np.float32(x)) for x in c)))
    print("max error: FP64 %.2e, FP32 Horner %.2e" % (e64, e32))
