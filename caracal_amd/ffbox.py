"""The QMDFF-form periodic box: system builder and parameters.

Functional form (see csrc/ff_box.hip): Morse O-H bonds, harmonic H-O-H
angles k (theta - theta0)^2, Zahn-damped Coulomb, BJ-damped D3-form C6/C8
dispersion and the QMDFF Pauli repulsion, nonbonded between molecules only,
10 angstrom cutoff, minimum image in a cubic box.

Stated parameter choices (the reference has no implementation of this exact
box; the forms follow src/ff_eg.f90 and src/ff_nonb.f90):
  Morse      De = 0.2 Eh, a = 1.2 / bohr, r0 = 0.96 angstrom
  angle      k = 0.15 Eh/rad^2, theta0 = 104.5 deg
  charges    O -0.8 e, H +0.4 e
  C6         O-O 12.8, O-H 5.6, H-H 3.1 Eh bohr^6 (D3-like magnitudes for
             two-coordinated O and one-coordinated H)
  BJ damping a1 = 0.45, a2 = 4.0 bohr, s8 = 2.7 and the D3 <r4>/<r2> ratios
             of H and O, all as src/setnonb.f90:47-73 states them (default-REAL
             literals, rounded to binary32)
  repulsion  zab = Z_i Z_j with Z(O) = 6, Z(H) = 1; alpha = 1.5 (O-O),
             1.9 (O-H), 2.3 (H-H) / bohr
  Coulomb    Zahn a = 0.2 / angstrom, rc = 10 angstrom (src/set_periodic.f90:93-97)
"""
import math

import numpy as np

from .units import BOHR_ANGSTROM

NPAR = 10
PARAM_NAMES = ["De", "a", "r0", "k_angle", "theta0", "q_O", "q_H", "C6_OO", "C6_OH", "C6_HH"]


def _f32(x):
    return float(np.float32(x))


R2R4 = {"H": _f32(2.00734898), "O": _f32(2.59361680)}
A1, A2, S8 = _f32(0.45), 4.0, _f32(2.7)


def theta_ref():
    return np.array([0.2, 1.2, 0.96 / BOHR_ANGSTROM, 0.15, math.radians(104.5), -0.8, 0.4, 12.8, 5.6, 3.1])


def pair_tables():
    """Fixed per-type-pair tables, type 0 = O, 1 = H."""
    el = ["O", "H"]
    sr42 = np.zeros((2, 2))
    r0 = np.zeros((2, 2))
    z = {"O": 6.0, "H": 1.0}
    zrep = np.zeros((2, 2))
    arep = np.array([[1.5, 1.9], [1.9, 2.3]])
    for a in range(2):
        for b in range(2):
            ra, rb = R2R4[el[a]], R2R4[el[b]]
            sr42[a, b] = 3.0 * S8 * ra * rb
            r0[a, b] = A1 * math.sqrt(3.0 * ra * rb) + A2
            zrep[a, b] = z[el[a]] * z[el[b]]
    return sr42, r0, zrep, arep


def zahn_params(rc_angstrom=10.0):
    a = 0.2 * BOHR_ANGSTROM                     # 0.2 / angstrom in 1/bohr
    rc = rc_angstrom / BOHR_ANGSTROM
    zp = math.erfc(a * rc) / rc ** 2 + 2 * a / math.sqrt(math.pi) * math.exp(-(a * rc) ** 2) / rc
    return a, rc, zp


def box_system(nmol=1000, box_angstrom=31.0, seed=2, theta=None):
    """Bent triatomics on a cubic lattice with random orientations.
    Returns dict(x (natoms,3) bohr, type, mol, bonds, angles, box)."""
    theta = theta_ref() if theta is None else theta
    rng = np.random.default_rng(seed)
    box = box_angstrom / BOHR_ANGSTROM
    m = int(math.ceil(nmol ** (1.0 / 3.0)))
    sites = np.array([(i, j, k) for i in range(m) for j in range(m) for k in range(m)], dtype=np.float64)[:nmol]
    sites = (sites + 0.5) * (box / m)
    r0, th0 = theta[2], theta[4]
    local = np.array([[0.0, 0.0, 0.0],
                      [r0, 0.0, 0.0],
                      [r0 * math.cos(th0), r0 * math.sin(th0), 0.0]])
    local -= local.mean(axis=0)
    x = np.zeros((3 * nmol, 3))
    for n in range(nmol):
        rot, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        x[3 * n:3 * n + 3] = sites[n] + local @ rot.T
    x = np.mod(x, box)
    types = np.tile(np.array([0, 1, 1], dtype=np.int32), nmol)
    mol = np.repeat(np.arange(nmol, dtype=np.int32), 3)
    bonds = np.array([[3 * n, 3 * n + 1] for n in range(nmol)] + [[3 * n, 3 * n + 2] for n in range(nmol)],
                     dtype=np.int32)
    angles = np.array([[3 * n + 1, 3 * n, 3 * n + 2] for n in range(nmol)], dtype=np.int32)
    return dict(x=x, type=types, mol=mol, bonds=bonds, angles=angles, box=box)


def masses(types):
    from .units import mass_au
    return np.where(np.asarray(types) == 0, mass_au(15.99491461956), mass_au(1.00782503207))
