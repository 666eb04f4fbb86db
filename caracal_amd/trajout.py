"""Trajectory output of the reference's MD step (src/verlet.f90:783-957), written
on the host from engine snapshots.  The engine keeps the state in HBM; a caller
that wants the reference's files runs the engine iwrite steps at a time and
hands each snapshot (get_state / get_forces) to these writers, which matches
the reference's mod(istep, iwrite) == 0 frames.

Files and their layout, as the reference writes them:

trajectory.xyz      natoms*nbeads / blank line (box lengths in angstrom under
                    NPT) / one line per bead and atom: name x y z (angstrom,
                    bead-major) (src/verlet.f90:789-799, opened by
                    src/dynamic.f90:1480)
md_gradients.dat    natoms*nbeads / blank / atom bead dV/dx dV/dy dV/dz
                    (hartree/bohr) (src/verlet.f90:939-944)
md_velocities.dat   natoms*nbeads / blank / atom bead vx vy vz (p/m, atomic
                    units) (src/verlet.f90:953-958)
traj_gen.xyz,       natoms / "Generation step No. istep" / name x y z of bead 1
traj_cross.xyz      (angstrom); with traj_gen_xi.dat / traj_cross_xi.dat: xi_real
                    xi_ideal per frame (src/verlet.f90:897-913,
                    src/calc_rate.f90:820-824 and 2211-2215)
MACE frames         natoms / Properties=species:S:1:pos:R:3:REF_forces:R:3
                    REF_energy=E pbc="F F F" / name, positions (angstrom, f14.8)
                    and forces (eV/angstrom, f14.8) of bead 1; E in eV
                    (src/verlet.f90:883-893)

The reference writes free-form records with list-directed output, whose
spacing depends on its compiler; these writers keep the records and their
fields and print every list-directed real with 17 significant digits, so a
frame reads back to the last bit.  The fixed formats (the MACE lines) are
reproduced exactly.  The VASP XDATCAR/CONTCAR frames (src/verlet.f90:806-880)
are not written.
"""
import os

import numpy as np

BOHR = 0.52917721092       # angstrom per bohr (src/general.f90:256)
EVOLT = 27.21138503        # eV per hartree (src/general.f90:258)


def _r(x):
    return f"{float(x):25.17E}"


def _frames(a, natoms):
    """(nbeads, natoms, 3) view of one trajectory's beads (a single (natoms, 3) frame is one bead)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 2:
        a = a[None]
    if a.shape[-1] != 3 or a.shape[-2] != natoms:
        raise ValueError(f"expected (nbeads, {natoms}, 3), got {a.shape}")
    return a


class TrajectoryWriter:
    """Writers of one trajectory's files.  names: element symbols of the atoms.
    Each write_* call appends one frame; flush() or close() ends the files."""

    def __init__(self, names, directory=".", *, verbose=False, box=None):
        self.names = list(names)
        self.natoms = len(self.names)
        self.dir = directory
        self.box = None if box is None else np.asarray(box, dtype=np.float64)
        self._files = {}
        self._open = lambda name: self._files.setdefault(name, open(os.path.join(self.dir, name), "w"))
        self.verbose = verbose

    # -- dynamic.x ---------------------------------------------------------------
    def write_positions(self, q):
        """One trajectory.xyz frame; q in bohr, (nbeads, natoms, 3)."""
        q = _frames(q, self.natoms)
        f = self._open("trajectory.xyz")
        f.write(f"{self.natoms * q.shape[0]:12d}\n")
        f.write("" if self.box is None else " ".join(_r(b * BOHR) for b in self.box))
        f.write("\n")
        for k in range(q.shape[0]):
            for i in range(self.natoms):
                f.write(f" {self.names[i]:<2s}" + "".join(_r(x * BOHR) for x in q[k, i]) + "\n")

    def _indexed(self, name, a):
        a = _frames(a, self.natoms)
        f = self._open(name)
        f.write(f"{self.natoms * a.shape[0]:12d}\n\n")
        for k in range(a.shape[0]):
            for i in range(self.natoms):
                f.write(f"{i + 1:12d}{k + 1:12d}" + "".join(_r(x) for x in a[k, i]) + "\n")

    def write_gradients(self, g):
        """One md_gradients.dat frame (verbose runs); g in hartree/bohr."""
        self._indexed("md_gradients.dat", g)

    def write_velocities(self, p, masses):
        """One md_velocities.dat frame (verbose runs): p / m in atomic units."""
        p = _frames(p, self.natoms)
        self._indexed("md_velocities.dat", p / np.asarray(masses, dtype=np.float64)[None, :, None])

    def write_md_frame(self, q, g=None, p=None, masses=None):
        """Everything the reference writes at a frame of dynamic.x."""
        self.write_positions(q)
        if self.verbose:
            self.write_gradients(g)
            self.write_velocities(p, masses)

    # -- calc_rate.x (print_gen / print_cross) -------------------------------------
    def write_rate_frame(self, q, istep, xi_real, xi_ideal, kind="gen"):
        """One traj_{gen,cross}.xyz frame (bead 1) and its xi line."""
        if kind not in ("gen", "cross"):
            raise ValueError("kind is 'gen' or 'cross'")
        q = _frames(q, self.natoms)
        f = self._open(f"traj_{kind}.xyz")
        f.write(f"{self.natoms:12d}\n Generation step No.{istep:12d}\n")
        for i in range(self.natoms):
            f.write(f" {self.names[i]:<2s}" + "".join(_r(x * BOHR) for x in q[0, i]) + "\n")
        self._open(f"traj_{kind}_xi.dat").write(_r(xi_real) + _r(xi_ideal) + "\n")

    def write_mace_frame(self, q, g, epot, xi=None, kind="gen"):
        """One MACE training-set frame of bead 1: energy in eV, forces -g in
        eV/angstrom (train_format MACE).  xi = (xi_real, xi_ideal) also
        writes the xi line, as under calc_rate."""
        q, g = _frames(q, self.natoms), _frames(g, self.natoms)
        f = self._open(f"traj_{kind}.xyz")
        f.write(f"{self.natoms:12d}\n")
        f.write(f"Properties=species:S:1:pos:R:3:REF_forces:R:3 REF_energy={epot * EVOLT:20.10f} pbc=\"F F F\"\n")
        for i in range(self.natoms):
            pos = "".join(f"{x * BOHR:14.8f}" for x in q[0, i])
            frc = "".join(f"{-x / BOHR * EVOLT:14.8f}" for x in g[0, i])
            f.write(f"{self.names[i]}  {pos} {frc}\n")
        if xi is not None:
            self._open(f"traj_{kind}_xi.dat").write(_r(xi[0]) + _r(xi[1]) + "\n")

    def flush(self):
        for f in self._files.values():
            f.flush()

    def close(self):
        for f in self._files.values():
            f.close()
        self._files.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_xyz_frames(path):
    """Frames of an xyz file written above: list of (comment, names, (n, 3) array)."""
    out = []
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines) and lines[i].strip():
        n = int(lines[i])
        comment = lines[i + 1]
        rows = [lines[i + 2 + j].split() for j in range(n)]
        out.append((comment, [r[0] for r in rows], np.array([[float(x) for x in r[1:4]] for r in rows])))
        i += 2 + n
    return out


def read_indexed_frames(path):
    """Frames of md_gradients.dat / md_velocities.dat: list of (n, 5) arrays
    (atom, bead, x, y, z)."""
    out = []
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines) and lines[i].strip():
        n = int(lines[i])
        out.append(np.array([[float(x) for x in lines[i + 2 + j].split()] for j in range(n)]))
        i += 2 + n
    return out
