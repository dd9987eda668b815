"""Coefficients of the FP32 arctangent used by the flux kernels (``atan2_pos_f``
in rr_device.hpp).

atan(t) = t * P(t^2) on t in [0, 1], P of degree 7, fitted here by iteratively
reweighted least squares on 4000 Chebyshev nodes (Lawson's algorithm for the
minimax polynomial: the weights grow where the error is largest).  Run this file
to print the coefficients and the error of the FP32 Horner evaluation; the
kernel's constants are these values rounded to FP32.  Max |error| over [0, 1]:
3.7e-8 for the polynomial, 1.4e-7 rad with FP32 arithmetic.
"""
import numpy as np


def fit(deg=7, nodes=4000, iters=60):
    t = np.sort(np.cos(np.pi * (np.arange(nodes) + 0.5) / nodes) * 0.5 + 0.5)
    s = t * t
    w = np.ones(nodes)
    V = np.vander(s, deg + 1, increasing=True)
    for _ in range(iters):
        c, *_ = np.linalg.lstsq(V * (t * w)[:, None], np.arctan(t) * w, rcond=None)
        err = np.abs(t * (V @ c) - np.arctan(t))
        w = w * (1.0 + err / err.max())
    return c


def fp32_error(c):
    t = np.linspace(0, 1, 200001).astype(np.float32)
    s = t * t
    cf = c.astype(np.float32)
    p = cf[-1]
    for k in range(len(cf) - 2, -1, -1):
        p = (p * s + cf[k]).astype(np.float32)
    return float(np.abs((t * p).astype(np.float64) - np.arctan(t.astype(np.float64))).max())


if __name__ == '__main__':
    c = fit()
    for k, x in enumerate(c):
        print(f'c{k} = {float(np.float32(x)):.9g}f')
    print('max |error| with FP32 evaluation:', fp32_error(c))
