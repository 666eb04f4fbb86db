"""Locates the first-order saddle of the BrH2 DIM-3C surface
(src/egrad_brh2.f) for H + HBr -> H2 + Br on the oracle: a root of the
gradient along the collinear Hc-Ha-Br line from a guess between the
reactant (Ha-Br at re = 2.673 bohr) and the product side, kept when the
central-difference Hessian has exactly one negative eigenvalue (and the five
zero modes of a linear molecule).  Prints the geometry that
caracal_amd/workloads.py holds as BRH2_TS (bohr; atoms Ha, Br, Hc; on the z
axis, centred on the mean position).
Run from the repository root: python tests/golden/make_brh2_saddle.py"""
import os
import sys

import numpy as np
from scipy.optimize import root

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
import oracle_lib as ol  # noqa: E402


def line(r):
    """Ha at the origin, Br at +r[0], Hc at -r[1] on z."""
    return np.array([[0, 0, 0], [0, 0, r[0]], [0, 0, -r[1]]], dtype=np.float64)


def grad(x):
    return ol.egrad_brh2(x.reshape(1, 3, 3))[1].ravel()


def hessian(x, h=1e-4):
    rows = []
    for i in range(9):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        rows.append((grad(xp) - grad(xm)) / (2 * h))
    hm = np.array(rows)
    return 0.5 * (hm + hm.T)


sol = root(lambda r: grad(line(r).ravel()).reshape(3, 3)[1:, 2], [2.8, 2.0], method="hybr",
           options=dict(xtol=1e-15))
x = line(sol.x)
x = x - x.mean(axis=0)
w = np.linalg.eigvalsh(hessian(x.ravel()))
assert (w < -1e-4).sum() == 1 and (np.abs(w) < 1e-6).sum() == 5, w
v = ol.egrad_brh2(x[None])[0][0]
print(f"R(Br-Ha) = {sol.x[0]:.12f}, R(Ha-Hc) = {sol.x[1]:.12f} bohr")
print(f"V = {v:.12f} hartree = {v * 627.5094743:.6f} kcal/mol, max |dV/dq| {np.abs(grad(x.ravel())).max():.1e}")
print(np.array2string(x, precision=16, separator=", "))
