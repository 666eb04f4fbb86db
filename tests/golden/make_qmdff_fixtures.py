"""Writes QMDFF fixtures: the reference's resolved parameters and its own
energies and gradients for a few structures.

Runs oracle/_ref/ref_qmdff (prepare + set_periodic + ff_eg/ff_nonb/ff_hb
compiled from the reference's sources) on

  ethanol      examples/poly_qmdff/ethanol/ethanol.qmdff (9 atoms, one molecule)
  glutathione  examples/dynamic/glutathione/glutathione.qmdff (37 atoms, 148
               listed hydrogen bonds, out-of-plane terms)
  dgevb        examples/evbopt/DG-EVB/min1.qmdff (6 atoms, 4 listed hydrogen bonds)
  ethanol16    the first 16 molecules of examples/dynamic/ethanol_box/box.qmdff
               (144 atoms) in the example's periodic 27 angstrom box:
               intermolecular dispersion, repulsion, Coulomb and hydrogen bonds

and stores, per system, qmdff_<name>.npz: the parameter arrays (element-pair
tables cut to elements 1..18, which hold every element used), three
structures (the file's, and two seeded perturbations of 0.1 and 0.3 bohr)
and the reference's running energies after ff_eg, ff_nonb, ff_hb and its
gradient.  Also qmdff_box_positions.npz: the coordinates and molecule numbers
of the whole ethanol box (1125 atoms) with its box and cutoffs, from which
the tests build the 125-molecule box from the ethanol molecule's parameters;
and qmdff_ethanol16_nhc_reference.npz: 20 steps of the reference's dynamic.x
verlet on the 16-molecule box with 4 beads, the Nose-Hoover chain at 200 K
(damping 20 dt), dt 0.5 fs, the periodic wrap and transrot, from seeded
positions and momenta (frames at steps 1, 10 and 20).
Run where /root/reference is mounted:
    python tests/golden/make_qmdff_fixtures.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import refqmdff  # noqa: E402
import refstep  # noqa: E402
from caracal_amd import qmdff  # noqa: E402
from caracal_amd.units import FS_TO_AU, beta_of, mass_au  # noqa: E402

AMU = {1: 1.008, 6: 12.011, 8: 15.999}

EX = "/root/reference/examples"
NZ = 18
KEYS = ("at", "molnum", "q", "bond", "vbond", "angl", "vangl", "tors", "vtors", "hb", "vhb", "nci", "c6xy",
        "eps1", "eps2", "rad", "scalehb", "coul_cut", "vdw_cut", "cut_low", "zahn_a", "zahn_par", "box",
        "periodic", "zahn", "nmols", "natoms")


def subset_file(src, nmol, path):
    """A QMDFF file of the first nmol molecules of a box file (same format)."""
    t = qmdff.read_qmdff(src)
    keep = t["molnum"] <= nmol
    n = int(keep.sum())
    lines = open(src).read().splitlines()
    out = [f"{n:5d} {float(nmol):9.4f} {0.0:9.4f}", lines[1]]
    out += lines[2:2 + n]
    sel = lambda idx: np.all(idx <= n, axis=1)
    b, a, tr, nc = sel(t["bond"]), sel(t["angl"]), sel(t["tors"][:, :4]), sel(t["nci"][:, :2])
    out.append(f"{int(b.sum()):10d}{int(a.sum()):10d}{int(tr.sum()):10d}{0:10d}{int(nc.sum()):10d}{0:10d}")
    # the term records of the file, in order, for the kept atoms
    pos = 2 + t["natoms"] + 1
    rec = lines[pos:pos + len(b)]
    out += [r for r, k in zip(rec, b) if k]
    pos += len(b)
    rec = lines[pos:pos + len(a)]
    out += [r for r, k in zip(rec, a) if k]
    pos += len(a)
    rec = lines[pos:pos + len(tr)]
    out += [r for r, k in zip(rec, tr) if k]
    trip = t["nci"][nc]
    for s in range(0, len(trip), 8):
        out.append("".join(f"{v:6d}" for row in trip[s:s + 8] for v in row))
    open(path, "w").write("\n".join(out) + "\n")
    return t["xyz"][:n]


def save(name, path, xyz, **kw):
    rng = np.random.default_rng(4)
    xs = np.stack([xyz, xyz + rng.normal(0, 0.1, xyz.shape), xyz + rng.normal(0, 0.3, xyz.shape)])
    e1, e2, e3, g, p = refqmdff.run(path, xs, dump=True, **kw)
    data = {k: np.asarray(p[k]) for k in KEYS}
    for k in ("r094", "sr42", "zab", "r0ab"):
        data[k + "_18"] = np.asarray(p[k])[:NZ, :NZ]
    assert int(np.max(p["at"])) <= NZ
    np.savez_compressed(os.path.join(HERE, f"qmdff_{name}.npz"), x=xs, e_eg=e1, e_nonb=e2, e=e3, g=g, **data)


def main():
    if not refqmdff.available():
        raise SystemExit("oracle/_ref/ref_qmdff is not built (make -C oracle -f Makefile.ref)")
    for name, rel in (("ethanol", "poly_qmdff/ethanol/ethanol.qmdff"),
                      ("glutathione", "dynamic/glutathione/glutathione.qmdff"),
                      ("dgevb", "evbopt/DG-EVB/min1.qmdff")):
        path = os.path.join(EX, rel)
        save(name, path, qmdff.read_qmdff(path)["xyz"])
    box = os.path.join(EX, "dynamic/ethanol_box/box.qmdff")
    with tempfile.TemporaryDirectory() as tmp:
        sub = os.path.join(tmp, "ethanol16.qmdff")
        xyz = subset_file(box, 16, sub)
        save("ethanol16", sub, xyz, periodic=True, box=27.0)
        nb, steps = 4, 20
        at = qmdff.read_qmdff(sub)["at"]
        m = mass_au([AMU[int(z)] for z in at])
        beta = float(beta_of(200.0))
        dt = 0.5 * FS_TO_AU
        rng = np.random.default_rng(9)
        q0 = xyz[None] + rng.normal(0, 0.02, (nb,) + xyz.shape)
        p0 = rng.normal(size=q0.shape) * np.sqrt(m[None, :, None] * nb / beta)
        fr, _ = refstep.run("qmdff", q0, p0, m, dt, beta, steps, constrain=-1, every=1, thermostat=2,
                            kelvin=1.0 / (0.316679e-5 * beta), nose_q=20 * dt,
                            qmdff=dict(file=sub, periodic=True, box=27.0))
        keep = np.array([1, 10, 20])
        np.savez_compressed(os.path.join(HERE, "qmdff_ethanol16_nhc_reference.npz"), q0=q0, p0=p0, masses=m, dt=dt,
                            beta=beta, nose_q=20 * dt, steps=keep, q=fr["q"][keep - 1], p=fr["p"][keep - 1],
                            epot=fr["epot"])
    t = qmdff.read_qmdff(box)
    _, _, _, _, p = refqmdff.run(box, t["xyz"][None], dump=True, periodic=True, box=27.0)
    np.savez_compressed(os.path.join(HERE, "qmdff_box_positions.npz"), xyz=t["xyz"], molnum=t["molnum"],
                        box=p["box"], coul_cut=p["coul_cut"], vdw_cut=p["vdw_cut"], cut_low=p["cut_low"],
                        zahn_a=p["zahn_a"], zahn_par=p["zahn_par"])


if __name__ == "__main__":
    main()
