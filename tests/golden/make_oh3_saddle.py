"""Locates the first-order saddle of the OH + H2 surface (src/egrad_oh3.f)
on the oracle: a least-squares root of the gradient from planar guesses,
kept when the central-difference Hessian has exactly one negative
eigenvalue.  Prints the geometry that caracal_amd/workloads.py holds as
OH3_TS (bohr; atoms O, H1, H2, H3; centred on the mean position).
Run from the repository root: python tests/golden/make_oh3_saddle.py"""
import os
import sys

import numpy as np
from scipy.optimize import least_squares

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
import oracle_lib as ol  # noqa: E402


def grad(x):
    return ol.egrad_oh3(x.reshape(1, 4, 3))[1].ravel()


def hessian(x, h=1e-4):
    rows = []
    for i in range(12):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        rows.append((grad(xp) - grad(xm)) / (2 * h))
    hm = np.array(rows)
    return 0.5 * (hm + hm.T)


best = None
for angle in np.linspace(90.0, 130.0, 5):
    th = np.radians(angle)
    x0 = np.array([[0, 0, 0], [1.83 * np.cos(th), 1.83 * np.sin(th), 0], [2.6, 0, 0], [4.15, 0.1, 0]]).ravel()
    x = least_squares(grad, x0, xtol=1e-15, ftol=1e-15, gtol=1e-15).x
    if (np.linalg.eigvalsh(hessian(x)) < -1e-4).sum() == 1:
        v = ol.egrad_oh3(x.reshape(1, 4, 3))[0][0]
        if best is None or v < best[0]:
            best = (v, x)
v, x = best
x = x.reshape(4, 3)
x = x - x.mean(axis=0)
x[:, 2] = 0.0                                  # planar
print(f"V = {v:.12f} hartree = {v * 627.5094743:.6f} kcal/mol, max |dV/dq| {np.abs(grad(x.ravel())).max():.1e}")
print(np.array2string(x, precision=16, separator=", "))
