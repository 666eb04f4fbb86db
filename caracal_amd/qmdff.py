"""The reference's QMDFF on the GPU: its file format and its module arrays.

read_qmdff(path) reads a QMDFF file as the reference's rdsolvff does
(src/rdsolvff.f90:36-202): the atoms with coordinates (bohr), charges and
molecule numbers, then the stretch, bend and torsion terms, the listed
hydrogen bonds and the noncovalent pair list.  Torsion phase shifts are
multiplied by pi on reading and listed hydrogen bonds get their strengths
from the charges (hbpara with the scalehb factors of setnonb), as the
reference does.

The nonbonded tables -- C6 of every atom pair, BJ radii, C8/C6 factors,
repulsion prefactors and exponents per element pair, covalent radii -- are
what the reference's prepare computes from its D3 reference data
(src/prepare.f90:87-130, src/setnonb.f90:47-73): the Fortran host passes
them as they sit in its modules.  qmdff_params() merges a read file with
such tables into the parameter dict ForceField.qmdff takes.
"""
import numpy as np

BOHR = 0.52917721092   # angstrom (src/set_periodic.f90, src/ff_nonb.f90)

# hydrogen- and halogen-bond scaling of setnonb (src/setnonb.f90:126-140)
SCALEHB = {7: 0.8, 8: 0.3, 9: 0.1, 15: 2.0, 16: 2.0, 17: 2.0, 34: 2.0, 35: 2.0}
SCALEXB = {17: 0.30, 35: 0.60, 53: 0.80, 85: 1.00}


def _sp(x):
    """A default-REAL (binary32) literal widened to double."""
    return float(np.float32(x))


def scalehb_table():
    t = np.zeros(94)
    for z, v in SCALEHB.items():
        t[z - 1] = _sp(v)
    return t


def scalexb_table():
    t = np.zeros(94)
    for z, v in SCALEXB.items():
        t[z - 1] = _sp(v)
    return t


def hbpara(a, b, q):
    """src/hbpara.f90"""
    e = np.exp(-a * q)
    return e / (e + b)


def read_qmdff(path):
    """The contents of a QMDFF file as rdsolvff reads them: dict with natoms,
    at, xyz (natoms, 3) bohr, q, molnum, nmols, bond (1-based, nbond x 2),
    vbond (x 3), angl (x 3, vertex first), vangl (x 2), tors (x 6), vtors
    (x 14, phases times pi), hb (x 3), vhb (x 2), nci (x 3)."""
    with open(path) as f:
        lines = f.read().splitlines()
    pos = 0
    head = lines[pos].split()
    pos += 1
    n = int(float(head[0]))
    pos += 1                                             # charge model line
    at = np.zeros(n, dtype=np.int32)
    xyz = np.zeros((n, 3))
    q = np.zeros(n)
    molnum = np.zeros(n)
    tokens = []
    # the atom records are read list-directed: at, x, y, z, q, molnum
    while len(tokens) < 6 * n:
        tokens += lines[pos].split()
        pos += 1
    for i in range(n):
        r = tokens[6 * i:6 * i + 6]
        at[i] = int(r[0])
        xyz[i] = [float(v) for v in r[1:4]]
        q[i] = float(r[4])
        molnum[i] = float(r[5])
    counts = lines[pos].split()
    pos += 1
    nbond, nangl, ntors, nhb, nnci = (int(v) for v in counts[:5])

    def records(count, width):
        nonlocal pos
        out = []
        while len(out) < count:
            out.append(lines[pos].split()[:width] if width else lines[pos].split())
            pos += 1
        return out

    b = records(nbond, 5)
    bond = np.array([[int(r[0]), int(r[1])] for r in b], dtype=np.int32).reshape(nbond, 2)
    vbond = np.array([[float(v) for v in r[2:5]] for r in b]).reshape(nbond, 3)
    a = records(nangl, 5)
    angl = np.array([[int(r[0]), int(r[1]), int(r[2])] for r in a], dtype=np.int32).reshape(nangl, 3)
    vangl = np.array([[float(v) for v in r[3:5]] for r in a]).reshape(nangl, 2)
    tors = np.zeros((ntors, 6), dtype=np.int32)
    vtors = np.zeros((ntors, 14))
    for m in range(ntors):
        r = lines[pos].split()
        pos += 1
        tors[m] = [int(v) for v in r[:6]]
        nt = tors[m, 4]
        vals = [float(v) for v in r[6:8 + 3 * nt]]
        vtors[m, :2] = vals[:2]
        for j in range(nt):
            vtors[m, 2 + 3 * j:5 + 3 * j] = vals[2 + 3 * j:5 + 3 * j]
            vtors[m, 3 + 3 * j] *= np.pi                    # src/rdsolvff.f90:105-109
    def triples(count):
        nonlocal pos
        vals = []
        while len(vals) < 3 * count:
            vals += [int(v) for v in lines[pos].split()]
            pos += 1
        return np.array(vals[:3 * count], dtype=np.int32).reshape(count, 3)
    hb = triples(nhb)
    sh = scalehb_table()
    sx = scalexb_table()
    vhb = np.zeros((nhb, 2))
    for k in range(nhb):                                 # src/rdsolvff.f90:140-152
        a1, a2, h = hb[k] - 1
        if at[h] == 1:
            vhb[k, 0] = hbpara(10.0, 5.0, q[a1]) * sh[at[a1] - 1]
            vhb[k, 1] = hbpara(10.0, 5.0, q[a2]) * sh[at[a2] - 1]
        else:
            vhb[k, 0] = sx[at[h] - 1] * hbpara(-6.5, 1.0, q[h])
    nci = triples(nnci)
    return dict(natoms=n, at=at, xyz=xyz, q=q, molnum=molnum, nmols=int(molnum.max()), bond=bond, vbond=vbond,
                angl=angl, vangl=vangl, tors=tors, vtors=vtors, hb=hb, vhb=vhb, nci=nci)


def qmdff_params(topology, tables):
    """A read file plus the nonbonded tables of prepare/set_periodic (c6xy,
    r094, sr42, zab, r0ab as [z_a - 1][z_b - 1], eps1, eps2, rad (bohr),
    scalehb, periodic, zahn, box, coul_cut, vdw_cut, cut_low, zahn_a,
    zahn_par) -> the parameter dict of ForceField.qmdff."""
    p = dict(topology)
    p.update(tables)
    return p
