"""Loads the QMDFF fixtures of tests/golden (written by
tests/golden/make_qmdff_fixtures.py from the compiled reference) as
parameter dicts in the layout of caracal_amd.qmdff.qmdff_params, and builds
the 125-molecule ethanol box from the ethanol molecule's parameters."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SYSTEMS = ("ethanol", "glutathione", "dgevb", "ethanol16")


def load(name):
    """(params, structures (3, natoms, 3), reference dict with e_eg, e_nonb, e, g)."""
    d = dict(np.load(os.path.join(GOLDEN, f"qmdff_{name}.npz")))
    p = {}
    for k, v in d.items():
        if k in ("x", "e_eg", "e_nonb", "e", "g"):
            continue
        if k.endswith("_18"):
            full = np.zeros((94, 94))
            full[:18, :18] = v
            p[k[:-3]] = full
        else:
            p[k] = v
    for k in ("natoms", "nmols", "periodic", "zahn"):
        p[k] = int(p[k])
    ref = {k: d[k] for k in ("e_eg", "e_nonb", "e", "g")}
    return p, d["x"], ref


def ethanol_box():
    """The example's ethanol box (125 molecules, 1125 atoms, 27 angstrom,
    periodic): every molecule carries the parameters of the ethanol molecule
    (examples/poly_qmdff/ethanol/ethanol.qmdff, whose terms the box file
    repeats molecule by molecule); C6 of an atom pair is that of the two atoms'
    places in the molecule.  Returns (params, positions (1125, 3))."""
    m, _, _ = load("ethanol")
    b = dict(np.load(os.path.join(GOLDEN, "qmdff_box_positions.npz")))
    na = m["natoms"]
    nmol = int(b["molnum"].max())
    n = na * nmol
    off = lambda a, k: a + k * na
    p = dict(m)
    p["natoms"], p["nmols"] = n, nmol
    p["at"] = np.tile(m["at"], nmol)
    p["q"] = np.tile(m["q"], nmol)
    p["molnum"] = b["molnum"].astype(np.float64)
    p["bond"] = np.concatenate([off(m["bond"], k) for k in range(nmol)])
    p["vbond"] = np.tile(m["vbond"], (nmol, 1))
    p["angl"] = np.concatenate([off(m["angl"], k) for k in range(nmol)])
    p["vangl"] = np.tile(m["vangl"], (nmol, 1))
    tors = []
    for k in range(nmol):
        t = m["tors"].copy()
        t[:, :4] += k * na
        tors.append(t)
    p["tors"] = np.concatenate(tors)
    p["vtors"] = np.tile(m["vtors"], (nmol, 1))
    nci = []
    for k in range(nmol):
        t = m["nci"].copy()
        t[:, :2] += k * na
        nci.append(t)
    p["nci"] = np.concatenate(nci)
    p["hb"] = np.zeros((0, 3), dtype=np.int32)
    p["vhb"] = np.zeros((0, 2))
    idx = np.arange(n) % na
    p["c6xy"] = m["c6xy"][np.ix_(idx, idx)]
    p["periodic"], p["zahn"] = 1, 0
    for k in ("box", "coul_cut", "vdw_cut", "cut_low", "zahn_a", "zahn_par"):
        p[k] = b[k]
    return p, b["xyz"]
