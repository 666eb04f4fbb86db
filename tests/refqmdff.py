"""Runs the reference's own QMDFF force field (oracle/_ref/ref_qmdff: prepare,
set_periodic, ff_eg, ff_nonb and ff_hb compiled from the reference's sources,
driven by oracle/ref_qmdff_harness.f90).  Test infrastructure only: pins the
oracle's QMDFF restatement and provides the parameters prepare resolves."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "oracle", "_ref", "ref_qmdff")


def available():
    return os.path.exists(BINARY)


class _Reader:
    def __init__(self, raw):
        self.raw, self.pos = raw, 0

    def take(self, dtype, count):
        a = np.frombuffer(self.raw, dtype=dtype, count=count, offset=self.pos)
        self.pos += a.nbytes
        return a.copy()


def run(qmdff_file, xyz, periodic=False, box=0.0, coul_cut=50.0, vdw_cut=10.0, zahn=False, dump=False):
    """Energies and gradients of structures xyz (nconf, natoms, 3) in bohr.
    box, coul_cut and vdw_cut in angstrom as in the key file.  Returns
    (e_eg, e_nonb, e_total) each (nconf,), g (nconf, natoms, 3) and, with
    dump, the resolved parameters as a dict."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nconf, n, _ = xyz.shape
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout, fdump = (os.path.join(tmp, f) for f in ("in.bin", "out.bin", "dump.bin"))
        with open(fin, "wb") as f:
            np.array([int(periodic), int(zahn), nconf], dtype=np.int32).tofile(f)
            np.array([box, coul_cut, vdw_cut], dtype=np.float64).tofile(f)
            xyz.tofile(f)
        args = [BINARY, os.path.abspath(qmdff_file), fin, fout] + ([fdump] if dump else [])
        subprocess.run(args, check=True, capture_output=True, text=True, cwd=tmp)
        out = np.fromfile(fout, dtype=np.float64).reshape(nconf, 3 + 3 * n)
        params = _parse_dump(open(fdump, "rb").read()) if dump else None
    e = out[:, :3]
    g = out[:, 3:].reshape(nconf, n, 3)
    return e[:, 0], e[:, 1], e[:, 2], g, params


def _parse_dump(raw):
    r = _Reader(raw)
    n, nbond, nangl, ntors, nhb, nnci, nmols = r.take(np.int32, 7)
    p = dict(natoms=int(n), nmols=int(nmols))
    p["at"] = r.take(np.int32, n)
    p["molnum"] = r.take(np.float64, n)
    p["q"] = r.take(np.float64, n)
    p["bond"] = r.take(np.int32, 2 * nbond).reshape(nbond, 2)
    p["vbond"] = r.take(np.float64, 3 * nbond).reshape(nbond, 3)
    p["angl"] = r.take(np.int32, 3 * nangl).reshape(nangl, 3)
    p["vangl"] = r.take(np.float64, 2 * nangl).reshape(nangl, 2)
    p["tors"] = r.take(np.int32, 6 * ntors).reshape(ntors, 6)
    p["vtors"] = r.take(np.float64, 14 * ntors).reshape(ntors, 14)
    p["hb"] = r.take(np.int32, 3 * nhb).reshape(nhb, 3)
    p["vhb"] = r.take(np.float64, 2 * nhb).reshape(nhb, 2)
    p["nci"] = r.take(np.int32, 3 * nnci).reshape(nnci, 3)
    p["c6xy"] = r.take(np.float64, n * n).reshape(n, n)      # symmetric
    for name in ("r094", "sr42", "zab", "r0ab"):
        p[name] = r.take(np.float64, 94 * 94).reshape(94, 94).T   # [z_i - 1][z_j - 1]
    p["eps1"] = r.take(np.float64, 6)
    p["eps2"] = r.take(np.float64, 6)
    p["rad"] = r.take(np.float64, 94)
    p["scalehb"] = r.take(np.float64, 94)
    p["scalexb"] = r.take(np.float64, 94)
    (p["coul_cut"], p["vdw_cut"], p["cut_low"], p["zahn_a"], p["zahn_par"],
     bx, by, bz) = r.take(np.float64, 8)
    p["box"] = np.array([bx, by, bz])
    p["periodic"] = int(r.take(np.int32, 1)[0])
    p["zahn"] = int(r.take(np.int32, 1)[0])
    return p
