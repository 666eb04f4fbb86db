"""Seeded test geometries for the analytic surfaces (bohr, numpy layout
(nbeads, natoms, 3) == Fortran q(3,natoms,nbeads))."""
import numpy as np

# Collinear H3 saddle of BKMP2: H-H = 1.757 bohr (examples/calc_rate/h+h2/ts.xyz).
H3_TS = np.array([[0.0, 0.0, -1.757], [0.0, 0.0, 0.0], [0.0, 0.0, 1.757]])


def h3_cloud(n, sigma, seed):
    rng = np.random.default_rng(seed)
    return H3_TS[None] + rng.normal(0.0, sigma, (n, 3, 3))


def h3_special():
    """Geometries that exercise every branch of BKMP2: compact (some R < 1.15,
    R < 0.95 Johnson branch), equilateral (cusp rounding), far dissociated."""
    out = []
    for r in (0.8, 0.9, 1.0, 1.1, 1.2, 1.4, 2.0, 3.0, 6.0):
        # equilateral triangle of side r
        out.append([[0, 0, 0], [r, 0, 0], [r / 2, r * np.sqrt(3) / 2, 0]])
        # collinear symmetric
        out.append([[0, 0, -r], [0, 0, 0], [0, 0, r]])
        # asymmetric bent
        out.append([[0, 0, 0], [1.4, 0, 0], [1.4 + r * 0.8, r * 0.6, 0]])
    out.append([[0, 0, 0], [1.4, 0, 0], [30.0, 0, 0]])
    return np.array(out, dtype=np.float64)


_A = 2.06  # C-H distance of the methane fragment, bohr
_TET = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=np.float64) / np.sqrt(3.0)


def ch4h_base(d_hb=2.0, stretch=1.0):
    """Atom order of the CH5 surface: 1 = h4, 2 = C, 3..5 = h1..h3, 6 = Hb
    (Hb on the far side of h1 at distance d_hb from it)."""
    x = np.zeros((6, 3))
    x[2] = _TET[0] * _A * stretch
    x[3] = _TET[1] * _A
    x[4] = _TET[2] * _A
    x[0] = _TET[3] * _A
    x[5] = _TET[0] * (_A * stretch + d_hb)
    return x


def ch4h_cloud(n, sigma, seed, d_hb=2.0):
    rng = np.random.default_rng(seed)
    return ch4h_base(d_hb)[None] + rng.normal(0.0, sigma, (n, 6, 3))


def water_box(nwater, box, seed, jitter=0.05):
    """nwater flexible waters (O, H, H) near r_OH = 1 angstrom and the H-H
    distance 1.633 angstrom, on a cubic lattice of side box (bohr) with random
    orientations, plus Gaussian jitter (bohr)."""
    rng = np.random.default_rng(seed)
    bohr = 0.52917721092
    r0, rhh = 1.0 / bohr, 1.633 / bohr
    half = np.arcsin(0.5 * rhh / r0)
    m = int(np.ceil(nwater ** (1.0 / 3.0)))
    a = box / m
    sites = np.array([(i, j, k) for i in range(m) for j in range(m) for k in range(m)], dtype=np.float64)
    sites = (sites[rng.permutation(len(sites))[:nwater]] + 0.5) * a
    out = []
    for s in sites:
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        h1 = r0 * np.array([np.sin(half), np.cos(half), 0.0])
        h2 = r0 * np.array([-np.sin(half), np.cos(half), 0.0])
        out += [s, s + q @ h1, s + q @ h2]
    return np.array(out) + rng.normal(0.0, jitter, (3 * nwater, 3))
