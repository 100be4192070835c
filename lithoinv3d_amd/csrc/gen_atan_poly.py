"""Derives the odd polynomial used by the FP32 atan2 in pair_math.h.

atan(a) ~= a * P(a^2) on a in [0, 1], P of degree ``deg`` in s = a^2.  The coefficients
minimise the maximum absolute error of atan(a) (a discrete Remez exchange on 4000
points, started from a least-squares fit at Chebyshev nodes).  Run:
    python gen_atan_poly.py [deg]
"""
import sys

import numpy as np


def fit(deg, iters=40):
    n = 4000
    a = 0.5 * (1 - np.cos(np.pi * (np.arange(n) + 0.5) / n))  # Chebyshev nodes in [0,1]
    s = a * a
    A = np.stack([a * s ** k for k in range(deg + 1)], 1)
    y = np.arctan(a)
    c = np.linalg.lstsq(A, y, rcond=None)[0]
    # Lawson iteration (IRLS towards the minimax solution)
    w = np.ones(n)
    for _ in range(iters):
        c = np.linalg.lstsq(A * np.sqrt(w)[:, None], y * np.sqrt(w), rcond=None)[0]
        e = np.abs(A @ c - y)
        w = w * e
        w /= w.sum()
    return c, np.abs(A @ c - y).max()


if __name__ == "__main__":
    deg = int(sys.argv[1]) if len(sys.argv) > 1 else 7
    c, err = fit(deg)
    print("max abs err", err)
    for k, v in enumerate(c):
        print(f"c{k} = {v:.9e}f")
