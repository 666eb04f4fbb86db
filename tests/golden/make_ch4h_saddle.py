"""Locates the first-order saddle of the CH5 surface that the CH4 + H
workloads start from, using the oracle surface (partitioned rational-function
Newton steps: uphill along the lowest non-rigid mode, downhill along the
others).  Prints the geometry that caracal_amd/workloads.py embeds as CH4H_TS.

    python tests/golden/make_ch4h_saddle.py
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
import geometries as geo  # noqa: E402
import oracle_lib as ol  # noqa: E402


def eg(x):
    v, g = ol.egrad_ch4h(x.reshape(1, 6, 3))
    return v[0], g.reshape(-1)


def hessian(x, h=1e-4):
    hm = np.zeros((18, 18))
    for i in range(18):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        hm[:, i] = (eg(xp)[1] - eg(xm)[1]) / (2 * h)
    return 0.5 * (hm + hm.T)


def find_saddle(x, tol=1e-10, maxit=60):
    for _ in range(maxit):
        v, g = eg(x)
        if np.linalg.norm(g) < tol:
            break
        w, u = np.linalg.eigh(hessian(x))
        gq = u.T @ g
        keep = np.argsort(np.abs(w))[6:]             # drop translations and rotations
        keep = keep[np.argsort(w[keep])]
        step = np.zeros(18)
        for j, k in enumerate(keep):
            lam = abs(w[k])
            step += (gq[k] / lam if j == 0 else -gq[k] / lam) * u[:, k]
        n = np.linalg.norm(step)
        x = x + step * min(1.0, 0.2 / n)
    return x


if __name__ == "__main__":
    x = find_saddle(geo.ch4h_base(d_hb=1.7, stretch=1.35).reshape(-1))
    np.set_printoptions(precision=17)
    print(repr(x.reshape(6, 3)))
